"""The oracle reproduces the fixtures written by vapor's own CPU solver, bit for
bit, and the reference's published whole-step baselines
(tests/SimulationTests.cpp:17-34, tolerance 2e-5 relative as there)."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import (VARIANTS, baselines, bits_equal, density_sums, load, options, oracle_face, oracle_project,
                     oracle_scalar, ptr, rel_close, result_row, variant_params)

STAGES = ["stages_n5.npz", "stages_n9.npz"]
STEPS = ["steps_n8_stir.npz", "steps_n9_box1.npz", "steps_n12_seed.npz", "steps_n16_omega18.npz"]


@pytest.mark.parametrize("name", STAGES)
def test_advection_matches_reference(name):
    g = load(name)
    n, box = int(g["n"]), float(g["box"])
    flow = (g["flow_u"], g["flow_v"], g["flow_w"])
    checked = 0
    for di, dt in enumerate(g["dts"]):
        for ci, cfl in enumerate(g["cfls"]):
            for method in (0, 1):
                for axis in (0, 1, 2):
                    got = oracle_face(n, box, axis, method, float(dt), float(cfl), flow, g["src_face"])
                    assert bits_equal(got, g[f"face_a{axis}_m{method}_d{di}_c{ci}"]), (axis, method, dt, cfl)
                    checked += 1
                got = oracle_scalar(n, box, method, float(dt), float(cfl), flow, g["src_cell"])
                assert bits_equal(got, g[f"cell_m{method}_d{di}_c{ci}"]), (method, dt, cfl)
                checked += 1
    assert checked == 32


@pytest.mark.parametrize("name", STAGES)
def test_samplers_divergence_gradient_match_reference(name):
    g = load(name)
    n, box = int(g["n"]), float(g["box"])
    lib = oracle.oracle_lib()
    for axis in (-1, 0, 1, 2):
        f = g["src_cell"] if axis < 0 else g["src_face"]
        f = np.ascontiguousarray(f)
        got = np.array([lib.orc_sample(n, box, axis, ptr(f), *map(float, p)) for p in g["sample_points"]],
                       np.float32)
        assert bits_equal(got, g[f"sample_a{axis}"]), axis
    u, v, w = (np.ascontiguousarray(g[k]) for k in ("flow_u", "flow_v", "flow_w"))
    div = np.zeros(n ** 3, np.float32)
    lib.orc_divergence(n, box, ptr(u), ptr(v), ptr(w), ptr(div))
    assert bits_equal(div, g["divergence"])
    gu, gv, gw = u.copy(), v.copy(), w.copy()
    lib.orc_apply_gradient(n, box, ptr(np.ascontiguousarray(g["grad_p"])), np.float32(0.02), ptr(gu), ptr(gv),
                           ptr(gw))
    assert bits_equal(gu, g["grad_u"]) and bits_equal(gv, g["grad_v"]) and bits_equal(gw, g["grad_w"])


@pytest.mark.parametrize("name", STAGES)
def test_projection_matches_reference(name):
    g = load(name)
    n, box = int(g["n"]), float(g["box"])
    flow = (g["flow_u"], g["flow_v"], g["flow_w"])
    for k, (floor, iters, fixed) in enumerate(g["proj_cases"]):
        o = options(dt=1 / 60, relaxation=1.7, max_iterations=int(iters), fixed_iterations=int(fixed),
                    diagonal_floor=float(floor))
        (u, v, w), p, r, _ = oracle_project(n, box, o, flow)
        assert bits_equal(u, g[f"proj{k}_u"]) and bits_equal(v, g[f"proj{k}_v"]) and bits_equal(w, g[f"proj{k}_w"])
        assert bits_equal(p, g[f"proj{k}_p"])
        np.testing.assert_array_equal(result_row(r), g[f"proj{k}_result"])


@pytest.mark.parametrize("name", STEPS)
def test_whole_steps_match_reference(name):
    import json

    g = load(name)
    n, box, frames = int(g["n"]), float(g["box"]), int(g["frames"])
    overrides = json.loads(str(g["params"]))
    for refl, flags in VARIANTS:
        p = variant_params(refl, flags, **overrides)
        sim = oracle.OracleSim(n, box, p)
        for _ in range(frames):
            sim.advance(1 / 60)
        tag = f"r{refl}f{flags}"
        for k, field in enumerate(("rho", "u", "v", "w")):
            assert bits_equal(sim.field(k), g[f"{tag}_{field}"]), (tag, field)
        np.testing.assert_array_equal(result_row(sim.last_projection()), g[f"{tag}_result"])
        d = sim.diagnostics()
        np.testing.assert_array_equal(np.array([d.kinetic_energy, d.max_velocity, d.max_divergence], np.float32),
                                      g[f"{tag}_diag"])


def test_published_baselines():
    doc = baselines()
    assert len(doc["rows"]) == 16
    for row in doc["rows"]:
        p = variant_params(row["reflect"], row["flags"], stir_strength=0.0)
        sim = oracle.OracleSim(row["n"], 4.5, p)
        for _ in range(6):
            sim.advance(1 / 60)
        d = sim.diagnostics()
        total, weighted = density_sums(sim.field(0))
        assert rel_close(d.kinetic_energy, row["energy"]), row
        assert rel_close(d.max_velocity, row["max_velocity"]), row
        assert rel_close(d.max_divergence, row["divergence"]), row
        assert rel_close(total, row["density_sum"]), row
        assert rel_close(weighted, row["weighted_density"]), row
