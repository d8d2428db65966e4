"""The workloads' own configurations against the oracle on the GPU.

- 64^3 reference case with its configured capped-Gaussian emitter: simple
  semi-Lagrangian step, emitter at (0.5, 0.5, 0.15) (z up), radius 0.08,
  strength 1, buoyancy 1, no decay, omega 1.7, 100 fixed sweeps: every field
  bit-identical to the oracle over 20 steps. The profile's e^v is a fixed
  sequence of IEEE float operations on both sides (profile_exp /
  orc_profile_exp), so device and host agree bit for bit; the oracle's is
  pinned against libm in test_oracle_extensions.py.
- The 256^3 mixed-precision scene (simple semi-Lagrangian, two Gaussian
  emitters, N(0, 0.05^2) face noise, FP32 sweeps inside FP64 iterative
  refinement to relative L2 1e-7, at most 5 corrections) at 32^3 and 48^3
  against the oracle, and at 256^3 itself one step against the all-FP64 run
  of the same state.
- Batches of uneven scenes drawn from {32, 48, 64, 96, 128}^3 (the batched
  workload's sizes, reflection MacCormack, the 128^3 workload's seeded scene
  family) against the oracle scene by scene: fixed sweeps (every grid up to
  64^3 bit-identical; 96^3 and 128^3 within FP32 rounding, because the
  projection's two FP64 means cancel to about 1e-12 there and the device's
  block tree rounds them differently from the oracle's sequential loop), and
  the relative-L2 1e-5 tolerance solve with per-scene masking (the oracle's
  sweep counts, projection by projection, and fields within 1e-4 of their
  scale: the gauge's FP64 sum in the check sweep is ordered differently).
- The staged backend's advection contract (AdvectionTests.cpp:159-177,
  AdvectionContext.cpp:15-67): aliasing, sizes, axis, dt and CFL checks.

The oracle runs in a thread pool (its C calls release the GIL), one scene per
thread."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from helpers import bits_equal, max_abs
from vapor_amd import abi, mac3d
from vapor_amd.train_warm_start import scene as family_scene

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def say(*args):
    print("[gpu workload]", *args, flush=True)


def scale_of(a):
    return max(1.0, float(np.max(np.abs(a))))


def run_oracles(sims, steps):
    """Advance every oracle simulation `steps` steps, in parallel threads;
    returns per simulation the projection records of each step."""
    def work(sim):
        records = []
        for _ in range(steps):
            sim.advance(DT)
            records.append((sim.last_projection(1), sim.last_projection(0)))
        return records

    with ThreadPoolExecutor(max_workers=min(8, len(sims))) as pool:
        return list(pool.map(work, sims))


# ---------------------------------------------------------------- 64^3, Gaussian

def reference_case_gaussian():
    p = mac3d.MacGridParameters3D(mac3d.SimpleSemiLagrangian)
    p.emitterCenterX, p.emitterCenterY, p.emitterCenterZ = 0.5, 0.5, 0.15
    p.emitterRadius = 0.08
    p.buoyancy = 1.0
    p.update(smoke_decay=0.0, source_strength=1.0, sor_omega=1.7, project_iterations=100, emitter_profile=1)
    return p


def test_reference_case_gaussian_profile_64_bitwise():
    p = reference_case_gaussian()
    sim = mac3d.GpuMacSimulation3D(64, 1.0, p)
    ref = oracle.OracleSim(64, 1.0, p)
    say("64^3 simple SL, capped-Gaussian emitter, 20 steps against the oracle")
    for step in range(20):
        sim.advance(DT)
        ref.advance(DT)
        if (step + 1) % 5 == 0:
            for name, got, want in zip(("rho", "u", "v", "w"), sim.fields(), ref.fields()):
                assert bits_equal(got, want), (step + 1, name, max_abs(got, want))
            say("step", step + 1, "bit-identical")
    rho = sim.fields()[0]
    # the profile, not the solid sphere: graded values below the cap
    graded = rho[(rho > 0.05) & (rho < 0.95)]
    assert graded.size > 100
    sim.close()


# ------------------------------------------------------- mixed-precision scene

def mixed_scene(seed=0):
    """The 256^3 mixed-precision configuration's scene (bench.py mixed_params)."""
    rng = np.random.default_rng(1000 + seed)
    p = mac3d.MacGridParameters3D(mac3d.SimpleSemiLagrangian)
    p.emitterCenterX, p.emitterCenterY, p.emitterCenterZ = 0.5, 0.5, 0.15
    p.emitterRadius = 0.08
    p.buoyancy = 1.0
    p.update(smoke_decay=0.0, source_strength=1.0, emitter_profile=1, emitter2_enabled=1,
             emitter2_center_x=rng.uniform(0.3, 0.7), emitter2_center_y=rng.uniform(0.3, 0.7),
             emitter2_center_z=rng.uniform(0.1, 0.2), emitter2_radius=rng.uniform(0.05, 0.1),
             sor_omega=1.8, solve_mode=1, tolerance=1e-7, max_sweeps=40000, check_interval=16, warm_start=1,
             refinement_steps=5)
    return p


@pytest.mark.parametrize("n", [32, 48])
def test_mixed_precision_scene_matches_oracle(n):
    p = mixed_scene()
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    ref = oracle.OracleSim(n, 1.0, p)
    sim.addVelocityNoise(0.05, 2026)
    ref.add_velocity_noise(0.05, 2026)
    for name, got, want in zip(("u", "v", "w"), sim.fields()[1:], ref.fields()[1:]):
        assert bits_equal(got, want), name  # the same noise bits
    say(f"{n}^3 mixed precision: two Gaussian emitters, noise, FP64 refinement to 1e-7, 3 steps")
    for step in range(3):
        sim.advance(DT)
        ref.advance(DT)
        r, q = sim.lastProjection(), ref.last_projection(0)
        say("step", step + 1, "sweeps gpu", r.iterations, "oracle", q.iterations, "corrections", r.refinement_steps,
            q.refinement_steps, "residual", r.final_residual_l2)
        assert r.converged and r.final_residual_l2 <= 1e-7
        assert 1 <= r.refinement_steps <= 5
        assert (r.iterations, r.refinement_steps) == (q.iterations, q.refinement_steps)
        for name, got, want in zip(("rho", "u", "v", "w"), sim.fields(), ref.fields()):
            err = max_abs(got, want)
            # the refinement's FP64 residual means are sums in another order
            assert err <= 1e-5 * scale_of(want), (step + 1, name, err)
    sim.close()


def test_mixed_precision_256_one_step_against_fp64():
    p = mixed_scene()
    sim = mac3d.GpuMacSimulation3D(256, 1.0, p)
    sim.addVelocityNoise(0.05, 2026)
    shadow = mac3d.ShadowRun3D(sim, 64)
    say("256^3 mixed precision, one step beside the all-FP64 run")
    sim.advance(DT)
    shadow.advance(DT)
    r, q = sim.lastProjection(), shadow.lastProjection()
    say("sweeps", r.iterations, "corrections", r.refinement_steps, "residual", r.final_residual_l2,
        "fp64 sweeps", q.iterations, "residual", q.final_residual_l2)
    assert r.converged and r.final_residual_l2 <= 1e-7
    assert 1 <= r.refinement_steps <= 5
    assert q.converged and q.final_residual_l2 <= 1e-7
    errors = shadow.compare(sim)
    for name, e in errors.items():
        say(name, {k: f"{v:.3e}" for k, v in e.items()})
        assert e["ref_max"] > 0.0, name
        # one step's error at the FP32 rounding level of each field; the
        # pressure also carries the two solves' different stopping iterates
        # within the 1e-7 tolerance
        loose = 10.0 if name == "p" else 1.0
        assert e["max_abs"] <= loose * 1e-4 * e["ref_max"], (name, e)
        assert e["rms"] <= loose * 1e-5 * e["ref_rms"], (name, e)
    shadow.close()
    sim.close()


# ------------------------------------------------------------------ batches

def batch_scene(k, fixed):
    p = family_scene(100 + k, mac3d)
    if fixed:  # an odd sweep count: half-steps of 7 sweeps exercise the ping-pong parity
        p.update(solve_mode=0, warm_start=0, project_iterations=15)
    return p


def compare_batch(sizes, fixed, steps):
    params = [batch_scene(k, fixed) for k in range(len(sizes))]
    batch = mac3d.MacBatch3D(sizes, [1.0] * len(sizes), params)
    refs = [oracle.OracleSim(n, 1.0, p) for n, p in zip(sizes, params)]
    say("batch", sizes, "fixed" if fixed else "tolerance 1e-5", steps, "steps; tuning",
        {k: v for k, v in batch.tuning().items() if not k.endswith("_ms")})
    got_records = []
    for _ in range(steps):
        batch.advance(DT)
        got_records.append([(batch.scene(s).lastProjection(1), batch.scene(s).lastProjection(0))
                            for s in range(len(sizes))])
    want_records = run_oracles(refs, steps)
    for s, n in enumerate(sizes):
        view = batch.scene(s)
        for step in range(steps):
            for k in (0, 1):
                r, q = got_records[step][s][k], want_records[s][step][k]
                assert r.iterations == q.iterations, (n, s, step, k, r.iterations, q.iterations)
                if not fixed:
                    assert r.converged and r.final_residual_l2 <= 1e-5, (n, s, step, k, r.final_residual_l2)
        worst = {}
        for name, got, want in zip(("rho", "u", "v", "w"), view.fields(), refs[s].fields()):
            err = max_abs(got, want)
            worst[name] = err
            if fixed and n <= 64:
                assert bits_equal(got, want), (n, s, name, err)
            else:
                # tolerance solves: the check sweep's FP64 sum of the iterate
                # (the zero-mean gauge) is a block tree, not the oracle's
                # sequential loop, and is not exact from 48^3 up
                assert err <= 1e-4 * scale_of(want), (n, s, name, err)
        say(f"scene {s} ({n}^3): sweeps", [got_records[t][s][1].iterations for t in range(steps)],
            "max |gpu - oracle|", {k: f"{v:.2e}" for k, v in worst.items()})
    batch.close()


def test_batch_uneven_scenes_fixed_sweeps_match_oracle():
    compare_batch([32, 48, 64, 96, 128, 48], fixed=True, steps=2)


def test_batch_uneven_scenes_tolerance_match_oracle():
    # 128^3 tolerance solves from rest take minutes on one host core, so the
    # largest scene here is 96^3; 128^3 meets the oracle in fixed mode above
    # and in test_gpu_configs.py's single-scene tolerance step.
    compare_batch([32, 48, 64, 96, 32, 64], fixed=False, steps=2)


# -------------------------------------------------- advection contract (HIP)

def test_staged_advection_contract_on_the_gpu():
    """AdvectionTests.cpp:159-177 against the HIP staged backend, called
    through the C ABI itself (no Python-side checks in between): every
    rejection is VGX_EINVAL (the reference's invalid_argument) and writes
    nothing. An empty span of the reference is a null pointer here; sizes are
    the backend grid's."""
    n = 3
    lib = abi.load_library()
    backend = mac3d.GpuMacBackend(n, 4.5)
    h = backend.handle
    cells, faces = n ** 3, (n + 1) * n * n
    flow = [np.zeros(faces, np.float32) for _ in range(3)]
    src = np.ones(faces, np.float32)
    out = np.full(faces, 7.0, np.float32)
    rho = np.ones(cells, np.float32)
    P = abi.float_ptr
    u, v, w = (P(f) for f in flow)
    SL, MC = abi.SEMI_LAGRANGIAN, abi.MACCORMACK

    def scalar(field, dt=0.1, cfl=3.0, method=SL, fu=u, fv=v, fw=w):
        return lib.vgx_backend_advect_scalar(h, method, dt, cfl, fu, fv, fw, field)

    def face(axis, source, output, dt=0.1, cfl=3.0, method=SL):
        return lib.vgx_backend_advect_face(h, axis, method, dt, cfl, u, v, w, source, output)

    rejected = {
        "empty scalar field": scalar(None),
        "face output is the source": face(0, P(src), P(src)),
        "face output is a tracer velocity": face(0, P(src), u),
        "axis 99": face(99, P(src), P(out)),
        "empty face source": face(0, None, P(out)),
        "dt = -1": scalar(P(rho), dt=-1.0),
        "dt = NaN": scalar(P(rho), dt=float("nan")),
        "dt = inf": scalar(P(rho), dt=float("inf")),
        "cfl = 0": scalar(P(rho), cfl=0.0, method=MC),
        "cfl < 0": scalar(P(rho), cfl=-1.0, method=MC),
        "cfl = NaN": scalar(P(rho), cfl=float("nan"), method=MC),
        "empty tracer x": scalar(P(rho), fu=None),
        "scalar aliases tracer z": lib.vgx_backend_advect_scalar(h, SL, 0.1, 3.0, u, v, P(rho), P(rho)),
        "face output overlaps tracer x": face(1, P(src), abi.float_ptr(flow[0][1:])),
    }
    for case, status in rejected.items():
        assert status == abi.VGX_EINVAL, (case, status)
        assert lib.vgx_last_error().decode(), case
    assert np.all(out == 7.0) and np.all(rho == 1.0) and np.all(src == 1.0)
    # a valid call still runs: stationary flow leaves the field unchanged
    assert face(0, P(src), P(out)) == abi.VGX_OK
    assert np.all(out == 1.0)
    assert scalar(P(rho), method=MC) == abi.VGX_OK
    assert np.all(rho == 1.0)
    backend.close()
