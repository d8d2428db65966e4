"""GPU parity: the HIP path against the reference's fixtures and the oracle.

Every float kernel keeps the reference's operation order and the library is
built without contraction, so results are expected to agree bit for bit; the
asserted bar is the north-star tolerance (advection and forces within 1e-5 of
the field scale), with the bitwise agreement reported alongside."""
import json

import numpy as np
import pytest

import oracle
from helpers import (VARIANTS, baselines, bits_equal, density_sums, load, max_abs, options, oracle_face,
                     oracle_project, oracle_scalar, rel_close, variant_params)
from vapor_amd import mac3d

pytestmark = pytest.mark.gpu

STAGES = ["stages_n5.npz", "stages_n9.npz"]
STEPS = ["steps_n8_stir.npz", "steps_n9_box1.npz", "steps_n12_seed.npz", "steps_n16_omega18.npz"]


def say(*args):
    print("[gpu]", *args, flush=True)


def close(got, want, rel=1e-5):
    want = np.asarray(want, np.float64)
    scale = max(1.0, float(np.max(np.abs(want)))) if want.size else 1.0
    return max_abs(got, want) <= rel * scale


@pytest.mark.parametrize("name", STAGES)
def test_staged_advection_matches_reference(name):
    g = load(name)
    n, box = int(g["n"]), float(g["box"])
    say("staged advection", name, "n", n)
    backend = mac3d.GpuMacBackend(n, box)
    adv = mac3d.GpuAdvection(backend)
    flow = tuple(np.ascontiguousarray(g[k]) for k in ("flow_u", "flow_v", "flow_w"))
    exact = total = 0
    for di, dt in enumerate(g["dts"]):
        for ci, cfl in enumerate(g["cfls"]):
            for method in (0, 1):
                for axis in (0, 1, 2):
                    out = np.zeros_like(g["src_face"])
                    adv.advectFace(flow, axis, np.ascontiguousarray(g["src_face"]), out, float(dt), method, float(cfl))
                    want = g[f"face_a{axis}_m{method}_d{di}_c{ci}"]
                    assert close(out, want), (axis, method, dt, cfl, max_abs(out, want))
                    exact += bits_equal(out, want)
                    total += 1
                field = np.ascontiguousarray(g["src_cell"]).copy()
                adv.advectScalar(flow, field, float(dt), method, float(cfl))
                want = g[f"cell_m{method}_d{di}_c{ci}"]
                assert close(field, want), (method, dt, cfl, max_abs(field, want))
                exact += bits_equal(field, want)
                total += 1
    say(f"bit-exact {exact}/{total}")
    assert exact == total


@pytest.mark.parametrize("name", STAGES)
def test_staged_projection_matches_reference(name):
    g = load(name)
    n, box = int(g["n"]), float(g["box"])
    say("staged projection", name)
    backend = mac3d.GpuMacBackend(n, box)
    solver = mac3d.GpuPressureSolver(backend)
    for k, (floor, iters, fixed) in enumerate(g["proj_cases"]):
        state = mac3d.MacGridState3D(n, box)
        state.velocityX()[:] = g["flow_u"]
        state.velocityY()[:] = g["flow_v"]
        state.velocityZ()[:] = g["flow_w"]
        o = mac3d.ProjectionOptions(dt=1 / 60, relaxation=1.7, max_iterations=int(iters), fixed_iterations=int(fixed),
                                    diagonal_floor=float(floor))
        r = solver.project(state, o)
        for f, key in ((state.velocityX(), "u"), (state.velocityY(), "v"), (state.velocityZ(), "w")):
            assert close(f, g[f"proj{k}_{key}"]), (k, key, max_abs(f, g[f"proj{k}_{key}"]))
        assert close(solver.pressure(), g[f"proj{k}_p"], 1e-4), (k, max_abs(solver.pressure(), g[f"proj{k}_p"]))
        row = g[f"proj{k}_result"]
        assert r.iterations == int(row[0]) and bool(r.converged) == bool(row[1]), (k, r.as_dict(), row)
        assert r.initial_residual == pytest.approx(row[2], rel=1e-6, abs=1e-9)
        assert r.final_residual == pytest.approx(row[3], rel=1e-3, abs=1e-7)
        assert r.divergence_before == pytest.approx(row[5], rel=1e-6)
        assert r.divergence_after == pytest.approx(row[6], rel=1e-3, abs=1e-6)
        assert r.removed_rhs_mean == pytest.approx(row[7], rel=1e-6, abs=1e-9)


@pytest.mark.parametrize("name", STEPS)
def test_resident_steps_match_reference(name):
    g = load(name)
    n, box, frames = int(g["n"]), float(g["box"]), int(g["frames"])
    overrides = json.loads(str(g["params"]))
    exact = 0
    for refl, flags in VARIANTS:
        say("resident", name, "reflect", refl, "flags", flags)
        p = mac3d.MacGridParameters3D()
        p.update(**variant_params(refl, flags, **overrides).as_dict())
        sim = mac3d.GpuMacSimulation3D(n, box, p)
        for _ in range(frames):
            sim.advance(1 / 60)
        tag = f"r{refl}f{flags}"
        fields = sim.fields()
        for k, field in enumerate(("rho", "u", "v", "w")):
            want = g[f"{tag}_{field}"]
            assert close(fields[k], want), (tag, field, max_abs(fields[k], want))
            exact += bits_equal(fields[k], want)
        d = sim.diagnostics()
        dw = g[f"{tag}_diag"]
        assert rel_close(d.kinetic_energy, dw[0]) and rel_close(d.max_velocity, dw[1])
        assert rel_close(d.max_divergence, dw[2], 1e-4)
        assert sim.downloadedBytes() == 4 * (n ** 3 + 3 * (n + 1) * n * n)  # only the explicit snapshot
        sim.close()
    say(f"bit-exact fields {exact}/{4 * len(VARIANTS)}")


def test_published_baselines_on_gpu():
    for row in baselines()["rows"]:
        say("baseline", row["n"], row["reflect"], row["flags"])
        p = mac3d.MacGridParameters3D()
        p.update(**variant_params(row["reflect"], row["flags"], stir_strength=0.0).as_dict())
        sim = mac3d.GpuMacSimulation3D(row["n"], 4.5, p)
        for _ in range(6):
            sim.advance(1 / 60)
        d = sim.diagnostics()
        total, weighted = density_sums(sim.state().density())
        assert rel_close(d.kinetic_energy, row["energy"]), row
        assert rel_close(d.max_velocity, row["max_velocity"]), row
        assert rel_close(d.max_divergence, row["divergence"]), row
        assert rel_close(total, row["density_sum"]), row
        assert rel_close(weighted, row["weighted_density"]), row
        sim.resetState()
        assert sim.kineticEnergy() == 0.0
        assert np.all(sim.state().density() == 0)
        sim.setBoxSize(3.0)
        assert sim.boxSize() == 3.0


def test_seeded_replay_and_independent_streams():
    # SimulationTests.cpp:44-75
    say("seeded replay")
    sims = []
    for seed in (12345, 12345, 54321):
        p = mac3d.MacGridParameters3D(emitterRadius=0.18, emitterCenterZ=0.3, projectIterations=16, randomSeed=seed)
        sims.append(mac3d.GpuMacSimulation3D(8, 4.5, p))
    first, second, other = sims
    for _ in range(6):
        first.advance(1 / 60)
        other.advance(1 / 60)
        second.advance(1 / 60)
    a, b, c = first.fields(), second.fields(), other.fields()
    assert all(bits_equal(x, y) for x, y in zip(a, b))
    assert not bits_equal(a[1], c[1])
    first.resetState()
    for _ in range(6):
        first.advance(1 / 60)
    assert all(bits_equal(x, y) for x, y in zip(first.fields(), b))


def test_resident_controls():
    # GraphicsSmokeTests.cpp:22-80 (resident part)
    say("resident controls")
    p = mac3d.MacGridParameters3D(emitterRadius=0.22, emitterCenterZ=0.3, stirStrength=0.5, projectIterations=20)
    sim = mac3d.GpuMacSimulation3D(9, 1.0, p)
    for _ in range(8):
        sim.advance(1 / 60)
    assert sim.downloadedBytes() == 0
    assert sim.uploadedBytes() == 0
    sim.resetState()
    sim.parameters().emitterKickMultiplier = 0
    sim.parameters().buoyancy = 0
    sim.parameters().stirStrength = 0
    sim.advance(1 / 60)
    rho, u, v, w = sim.fields()
    assert np.all(np.abs(u) < 1e-5) and np.all(np.abs(v) < 1e-5) and np.all(np.abs(w) < 1e-5)
    assert rho.max() > 0
    sim.setBoxSize(2.0)
    assert np.all(sim.state().density() == 0)
    sim.parameters().sourceStrength = 0
    sim.advance(1 / 60)
    rho, u, v, w = sim.fields()
    assert np.all(u == 0) and np.all(v == 0) and np.all(w == 0)
    with pytest.raises(ValueError):
        sim.advance(0.0)
    sim.parameters().emitterKickMultiplier = -1
    with pytest.raises(ValueError):
        sim.advance(1 / 60)


def test_nonfinite_state_is_rejected():
    say("non-finite")
    sim = mac3d.GpuMacSimulation3D(8, 1.0, mac3d.MacGridParameters3D(sourceStrength=0.0))
    s = mac3d.MacGridState3D(8, 1.0)
    s.velocityX()[100] = np.nan
    sim.upload(s)
    with pytest.raises(ValueError):
        sim.advance(1 / 60)


@pytest.mark.parametrize("n,refl,flags", [(32, 1, 3), (48, 0, 2), (64, 1, 3)])
def test_larger_grids_match_oracle(n, refl, flags):
    say("larger grid", n, refl, flags)
    p = mac3d.MacGridParameters3D(emitterRadius=0.1, emitterCenterZ=0.15, stirStrength=0.5, projectIterations=24,
                                  sorOmega=1.8, reflection=refl, macCormackSmoke=flags & 1,
                                  macCormackVel=(flags >> 1) & 1)
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    ref = oracle.OracleSim(n, 1.0, p)
    for _ in range(3):
        sim.advance(1 / 60)
        ref.advance(1 / 60)
    for got, want in zip(sim.fields(), ref.fields()):
        assert close(got, want), max_abs(got, want)


def test_determinism_bitwise():
    say("determinism")
    outs = []
    for _ in range(2):
        p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
        p.emitterRadius = 0.1
        p.projectIterations = 30
        sim = mac3d.GpuMacSimulation3D(40, 1.0, p)
        for _ in range(3):
            sim.advance(1 / 60)
        outs.append(sim.fields() + (sim.pressure(),))
    assert all(bits_equal(a, b) for a, b in zip(*outs))


def test_graph_replayed_chunks_match_stream_launches():
    """Tolerance chunks replayed from captured HIP graphs (the default) give
    the same bits and sweep counts as the same launches on the stream, over
    steps that replay each graph many times with other work in between, and
    with timing on (which launches some chunks on the stream)."""
    say("graph replay vs stream launches")
    runs = []
    for mode, timing in ((1, False), (0, False), (0, True)):
        p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
        p.update(emitter_radius=0.12, sor_omega=1.8, solve_mode=1, tolerance=1e-5, max_sweeps=6000,
                 check_interval=8, warm_start=1, launch_mode=mode)
        sim = mac3d.GpuMacSimulation3D(24, 1.0, p)
        sim.set_timing(timing)
        iters = []
        for _ in range(5):
            sim.advance(1 / 60)
            iters.append((sim.lastProjection(1).iterations, sim.lastProjection(0).iterations))
        runs.append((iters, sim.fields() + (sim.pressure(),)))
    (it0, f0) = runs[0]
    assert sum(a + b for a, b in it0) > 5 * 16  # the solves ran several chunks
    for it, f in runs[1:]:
        assert it == it0
        assert all(bits_equal(a, b) for a, b in zip(f, f0))


def test_velocity_noise_then_steps_match_oracle():
    """Initial N(0, 0.05) face noise (counter-based, extension) is the
    oracle's bit for bit, and so are fixed-sweep steps from it, with both
    emitters on."""
    say("velocity noise")
    n = 20
    p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
    p.update(emitter_radius=0.15, emitter2_enabled=1, emitter2_radius=0.1, emitter2_center_x=0.7,
             emitter2_center_y=0.3, emitter2_center_z=0.2, project_iterations=24)
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    ref = oracle.OracleSim(n, 1.0, p)
    sim.addVelocityNoise(0.05, 2026)
    ref.add_velocity_noise(0.05, 2026)
    for got, want in zip(sim.fields(), ref.fields()):
        assert bits_equal(got, want)
    for _ in range(2):
        sim.advance(1 / 60)
        ref.advance(1 / 60)
    for got, want in zip(sim.fields(), ref.fields()):
        assert bits_equal(got, want), max_abs(got, want)
    with pytest.raises(ValueError):
        sim.addVelocityNoise(float("nan"), 1)


def test_tolerance_mode_matches_oracle():
    say("tolerance mode")
    n = 16
    p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
    p.update(emitter_radius=0.1, sor_omega=1.8, solve_mode=1, tolerance=1e-5, max_sweeps=4000, check_interval=8,
             warm_start=1)
    sim = mac3d.GpuMacSimulation3D(n, 1.0, p)
    ref = oracle.OracleSim(n, 1.0, p)
    for _ in range(4):
        sim.advance(1 / 60)
        ref.advance(1 / 60)
        r, q = sim.lastProjection(), ref.last_projection()
        assert r.final_residual_l2 <= 1e-5 * 1.01 and r.converged
        say("sweeps gpu", r.iterations, "oracle", q.iterations)
        # the chunked checks evaluate the same iterates as the oracle's
        assert r.iterations == q.iterations, (r.iterations, q.iterations)
    for got, want in zip(sim.fields(), ref.fields()):
        assert close(got, want, 1e-4), max_abs(got, want)


def test_128_step_properties():
    say("128^3 reflection MacCormack properties")
    p = mac3d.MacGridParameters3D(mac3d.ReflectionMacCormack)
    p.update(emitter_radius=0.08, emitter_center_z=0.15, sor_omega=1.8, project_iterations=100)
    sim = mac3d.GpuMacSimulation3D(128, 1.0, p)
    for _ in range(3):
        sim.advance(1 / 60)
    rho, u, v, w = sim.fields()
    assert all(np.all(np.isfinite(f)) for f in (rho, u, v, w))
    assert rho.min() >= 0.0 and rho.max() <= 1.00001
    r = sim.lastProjection()
    assert r.divergence_after < r.divergence_before
    U = u.reshape(128, 128, 129)
    assert np.all(U[:, :, 0] == 0) and np.all(U[:, :, 128] == 0)


@pytest.mark.parametrize("n,tol,refine", [(16, 1e-5, 0), (16, 1e-9, 5), (24, 1e-8, 4)])
def test_staged_l2_solve_and_refinement_match_oracle(n, tol, refine):
    say("staged L2 solve", n, tol, refine)
    rng = np.random.default_rng(n)
    flow = [rng.standard_normal((n + 1) * n * n).astype(np.float32) * 0.1 for _ in range(3)]
    o = mac3d.ProjectionOptions(use_l2=1, l2_tolerance=tol, max_iterations=6000, residual_check_interval=16,
                                relaxation=1.8, refinement_steps=refine, fixed_iterations=0)
    backend = mac3d.GpuMacBackend(n, 1.0)
    solver = mac3d.GpuPressureSolver(backend)
    state = mac3d.MacGridState3D(n, 1.0)
    for dst, src in zip((state.velocityX(), state.velocityY(), state.velocityZ()), flow):
        dst[:] = src
    r = solver.project(state, o)
    (u, v, w), p, q, _ = oracle_project(n, 1.0, options(use_l2=1, l2_tolerance=tol, max_iterations=6000,
                                                        residual_check_interval=16, relaxation=1.8,
                                                        refinement_steps=refine), flow)
    assert q.converged and r.converged, (r.as_dict(), q.as_dict())
    assert r.final_residual_l2 <= tol
    assert r.refinement_steps == q.refinement_steps
    say("sweeps gpu", r.iterations, "oracle", q.iterations)
    assert r.iterations == q.iterations, (r.iterations, q.iterations)
    scale = max(1e-30, float(np.max(np.abs(p))))
    assert max_abs(solver.pressure(), p) <= 1e-4 * scale
