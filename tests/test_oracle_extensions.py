"""Extensions with no reference code, pinned by closed forms and invariants:
the capped-Gaussian and second emitters, the LCG stream, the relative L2
tolerance stop, warm starts and FP64 iterative refinement."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import bits_equal, f32, options, oracle_project, ptr, sizes, variant_params
from vapor_amd.abi import VgxProjectionResult


def lcg(state, count):
    out = []
    for _ in range(count):
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        out.append(np.float32(state >> 8) * np.float32(1.0 / 16777216.0))
    return state, out


def emit(n, box, params, dt=1 / 60, seed=None):
    cells, faces = sizes(n)
    rho = np.zeros(cells, np.float32)
    u, v, w = (np.zeros(faces, np.float32) for _ in range(3))
    state = oracle.oracle_lib().orc_emit(n, box, ctypes.byref(params), dt, params.random_seed if seed is None else seed,
                                         ptr(rho), ptr(u), ptr(v), ptr(w))
    return state, rho, u, v, w


def test_emitter_consumes_one_lcg_draw_per_cell_in_sphere():
    # MacGridFluidSolver3D.cpp:33-36,140-165: one draw per in-sphere cell.
    n, box = 16, 1.0
    p = variant_params(0, 0, stir_strength=0.5, emitter_radius=0.2, emitter_center_z=0.4)
    state, rho, u, v, w = emit(n, box, p)
    inside = int(np.count_nonzero(rho))
    expected, draws = lcg(p.random_seed, inside)
    assert state == expected
    # Stir values are bounded by stir * dt * 60 / 2 per face contribution, two per face.
    assert np.max(np.abs(u)) <= 0.5 * 0.5 * 2 + 1e-6


def test_gaussian_profile_closed_form():
    n, box = 17, 1.0
    p = variant_params(0, 0, stir_strength=0.0, emitter_radius=0.25, emitter_center_x=0.5, emitter_center_y=0.5,
                       emitter_center_z=0.5, emitter_profile=1, source_strength=0.2)
    _, rho, *_ = emit(n, box, p)
    # Centre cell (8,8,8) sits at the sphere centre: q = 0 -> min(1, 3 s) = 0.6.
    assert rho[8 + n * (8 + n * 8)] == pytest.approx(0.6, rel=1e-6)
    assert 0.0 <= rho.min() and rho.max() <= 1.0
    h = box / n
    for (x, y, z) in [(10, 8, 8), (8, 11, 9), (6, 7, 10)]:
        d2 = (((x + .5) * h - .5) ** 2 + ((y + .5) * h - .5) ** 2 + ((z + .5) * h - .5) ** 2) / (.25 * box) ** 2
        assert rho[x + n * (y + n * z)] == pytest.approx(min(1, 3 * 0.2 * np.exp(-3.5 * d2)), rel=1e-5)
    p.source_strength = 4.0  # capped at 1 everywhere near the centre
    _, rho, *_ = emit(n, box, p)
    assert rho.max() == 1.0


def test_profile_exp_against_libm():
    """The Gaussian profile's e^v (fixed IEEE float operations, the same on the
    device) against libm's double exp over the profile's whole argument range
    [-3.5, 0]: every 61st float there, plus the range ends and the points where
    the reduction's k changes (v = (k - 1/2) ln 2)."""
    lib = oracle.oracle_lib()
    top = np.float32(3.5).view(np.uint32)
    mags = np.arange(0, int(top) + 1, 61, dtype=np.uint32).view(np.float32)
    ks = np.arange(-5, 1, dtype=np.float64)
    edges = np.float32((ks - 0.5) * np.log(2.0))
    around = np.concatenate([np.nextafter(edges, np.float32(-np.inf)), edges, np.nextafter(edges, np.float32(0))])
    vs = np.concatenate([-mags, around, np.float32([0.0, -3.5])]).astype(np.float32)
    got = np.array([lib.orc_profile_exp(float(v)) for v in vs[:: max(1, vs.size // 200000)]], np.float64)
    vs = vs[:: max(1, vs.size // 200000)]
    want = np.exp(vs.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    err = np.abs(got - want) / ulp
    assert err.max() <= 1.5, (float(err.max()), float(vs[np.argmax(err)]))
    assert lib.orc_profile_exp(0.0) == 1.0
    # monotone over the sample (as e^v must be), so the capped profile is too
    order = np.argsort(vs)
    assert np.all(np.diff(got[order]) >= 0.0)


def test_second_emitter_equals_two_sequential_emits():
    n, box = 20, 1.0
    p = variant_params(0, 0, stir_strength=0.5, emitter_radius=0.12, emitter_center_x=0.3, emitter_center_z=0.3,
                       emitter2_enabled=1, emitter2_radius=0.1, emitter2_center_x=0.7, emitter2_center_y=0.6,
                       emitter2_center_z=0.25)
    state, rho, u, v, w = emit(n, box, p)
    a = variant_params(0, 0, stir_strength=0.5, emitter_radius=0.12, emitter_center_x=0.3, emitter_center_z=0.3)
    b = variant_params(0, 0, stir_strength=0.5, emitter_radius=0.1, emitter_center_x=0.7, emitter_center_y=0.6,
                       emitter_center_z=0.25)
    cells, faces = sizes(n)
    r2 = np.zeros(cells, np.float32)
    u2, v2, w2 = (np.zeros(faces, np.float32) for _ in range(3))
    lib = oracle.oracle_lib()
    s1 = lib.orc_emit(n, box, ctypes.byref(a), 1 / 60, p.random_seed, ptr(r2), ptr(u2), ptr(v2), ptr(w2))
    s2 = lib.orc_emit(n, box, ctypes.byref(b), 1 / 60, s1, ptr(r2), ptr(u2), ptr(v2), ptr(w2))
    assert s2 == state
    for x, y in ((rho, r2), (u, u2), (v, v2), (w, w2)):
        assert bits_equal(x, y)


def flow_with_divergence(n, seed=5):
    rng = np.random.default_rng(seed)
    cells, faces = sizes(n)
    return [rng.standard_normal(faces).astype(np.float32) * 0.1 for _ in range(3)]


def test_l2_tolerance_stop_and_residual():
    n = 12
    flow = flow_with_divergence(n)
    o = options(use_l2=1, l2_tolerance=1e-5, max_iterations=5000, residual_check_interval=4, relaxation=1.8)
    (u, v, w), p, r, b = oracle_project(n, 1.0, o, flow)
    assert r.converged and r.final_residual_l2 <= 1e-5 and r.initial_residual_l2 == pytest.approx(1.0)
    # Stopped at the first check below the tolerance: 4 sweeps earlier it was above.
    o2 = options(use_l2=1, l2_tolerance=1e-5, max_iterations=r.iterations - 4, fixed_iterations=0,
                 residual_check_interval=4, relaxation=1.8)
    _, _, r2, _ = oracle_project(n, 1.0, o2, flow)
    assert r2.final_residual_l2 > 1e-5
    # The L2 residual equals a dense evaluation.
    bb = b.astype(np.float64).reshape(n, n, n)
    x = np.pad(p.astype(np.float64).reshape(n, n, n), 1, mode="edge")
    lap = 6 * x[1:-1, 1:-1, 1:-1] - (x[:-2, 1:-1, 1:-1] + x[2:, 1:-1, 1:-1] + x[1:-1, :-2, 1:-1] + x[1:-1, 2:, 1:-1]
                                     + x[1:-1, 1:-1, :-2] + x[1:-1, 1:-1, 2:])
    dense = np.linalg.norm(bb - lap) / np.linalg.norm(bb)
    assert dense == pytest.approx(r.final_residual_l2, rel=1e-6)


def test_warm_start_from_the_solution_needs_no_sweeps():
    n = 10
    flow = flow_with_divergence(n, 7)
    o = options(use_l2=1, l2_tolerance=1e-6, max_iterations=8000, residual_check_interval=8, relaxation=1.8)
    _, p, r, _ = oracle_project(n, 1.0, o, flow)
    assert r.converged and r.iterations > 0
    o.warm_start = 1
    _, p2, r2, _ = oracle_project(n, 1.0, o, flow, pressure=p)
    assert r2.iterations == 0 and r2.warm_started == 1 and r2.converged


@pytest.mark.parametrize("n", [8, 16])
def test_fp64_refinement_beats_the_fp32_floor(n):
    flow = flow_with_divergence(n, 9)
    # FP32 SOR alone stalls above 1e-9 relative L2 ...
    plain = options(use_l2=1, l2_tolerance=1e-9, max_iterations=6000, residual_check_interval=16, relaxation=1.8)
    _, _, rp, _ = oracle_project(n, 1.0, plain, flow)
    assert not rp.converged and rp.final_residual_l2 > 1e-9
    # ... FP64 refinement around it gets there in a few corrections.
    ir = options(use_l2=1, l2_tolerance=1e-9, max_iterations=6000, residual_check_interval=16, relaxation=1.8,
                 refinement_steps=5)
    (u, v, w), p, r, _ = oracle_project(n, 1.0, ir, flow)
    assert r.converged and r.final_residual_l2 <= 1e-9
    assert 1 <= r.refinement_steps <= 5
    assert r.divergence_after < rp.divergence_after
