"""GPU parity of the MAC 2D solvers (include/vapor_mac2d.h) through the C ABI.

The bar: bit-exact against the oracle (oracle/mac2d_oracle.c in its default
mode, whose e^v and x^y are the same IEEE sequences the GPU evaluates) for
every stage and trajectory; bit-exact against vapor's own fixtures where no
transcendental enters (advection, projection); within 2e-5 of the field scale
of vapor's fixtures where the emitter's e^v and the buoyancy's x^(1/4) enter
(those two differ from the C library's by at most one unit in the last
place). Sizes cover every code path: the register-tile projection for
T = 1..4 (n <= 128), the global-memory projection (n > 128), the historical
solver's tiles up to T = 6 and its global path, and the limiter in LDS and in
global memory. The reference's own 2D tests (Mac2DTests.cpp,
Historical2DTests.cpp) are mirrored on the GPU path."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from oracle import mac2d as O
from vapor_amd import mac2d
from vapor_amd.abi import VgxProjectionOptions, VgxProjectionResult

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
DT = 1.0 / 60.0
RESULT_FIELDS = ("iterations", "converged", "initial_residual", "final_residual", "residual_available",
                 "divergence_before", "divergence_after", "removed_rhs_mean")


def say(*args):
    print("[gpu 2d]", *args, flush=True)


def row(r):
    return np.array([getattr(r, f) for f in RESULT_FIELDS], np.float64)


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, a.dtype)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def close(got, want, rel=2e-5):
    want = np.asarray(want, np.float64)
    scale = max(1.0, float(np.max(np.abs(want))))
    return float(np.max(np.abs(np.asarray(got, np.float64) - want))) <= rel * scale


def options(dt, relaxation, iters, fixed=0, check=8, abs_tol=1e-6, rel_tol=1e-5):
    o = VgxProjectionOptions()
    o.dt, o.relaxation, o.max_iterations, o.residual_check_interval = dt, relaxation, iters, check
    o.absolute_tolerance, o.relative_tolerance, o.fixed_iterations = abs_tol, rel_tol, fixed
    return o


def gpu_project(n, box, o, u, v):
    state = mac2d.MacGridState2D(n, box)
    state.velocityX()[:] = u
    state.velocityY()[:] = v
    solver = mac2d.GpuPressureSolver2D()
    r = solver.project(state, o)
    return state.velocityX().copy(), state.velocityY().copy(), solver.pressure().copy(), r


def oracle_project(n, box, o, u, v):
    uu, vv = u.copy(), v.copy()
    p = np.zeros(n * n, np.float64)
    r = VgxProjectionResult()
    rc = O.oracle2d_lib().o2_project(n, box, ctypes.byref(o), O.fptr(uu), O.fptr(vv), O.dptr(p), ctypes.byref(r))
    return rc, uu, vv, p, r


# ------------------------------------------------------------------- stages

def test_advection_matches_reference_fixtures_bitwise():
    g = np.load(GOLDEN / "mac2d_stages.npz")
    n, box = 9, 1.0
    sl = mac2d.SemiLagrangian2D()
    for field in (0, 1, 2):
        src = g[f"adv_src{field}"]
        for k, dt in enumerate(g["adv_dts"]):
            out = np.empty_like(src)
            sl.advect(n, box, g["adv_u"], g["adv_v"], src, out, float(dt), field)
            assert bits_equal(out, g[f"adv_out{field}_{k}"]), (field, dt)
    say("advection: 3 fields x 4 dt bit-identical to vapor's fixtures")


@pytest.mark.parametrize("n", [1, 2, 17, 64, 130])
def test_advection_matches_oracle_bitwise(n):
    rng = np.random.default_rng(n)
    N, F = n * n, n * (n + 1)
    u = (rng.standard_normal(F) * 2 / n).astype(np.float32)
    v = (rng.standard_normal(F) * 2 / n).astype(np.float32)
    sl = mac2d.SemiLagrangian2D()
    for field, count in ((0, N), (1, F), (2, F)):
        src = rng.random(count).astype(np.float32)
        for dt in (0.0, 0.3, 7.0):
            got, want = np.empty_like(src), np.empty_like(src)
            sl.advect(n, 1.5, u, v, src, got, dt, field)
            assert O.oracle2d_lib().o2_advect(n, 1.5, field, dt, O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(want)) == 0
            assert bits_equal(got, want), (n, field, dt)
    say("advection n", n, "bit-identical to the oracle")


def test_advection_contract_on_the_gpu():
    """SemiLagrangian2D.cpp:30-37 / Mac2DTests.cpp:83-91: bad dt, aliasing,
    wrong field, non-finite input and backtrace are rejected."""
    n = 8
    F = n * (n + 1)
    u = np.full(F, .5, np.float32)
    v = np.full(F, -.25, np.float32)
    src = np.ones(n * n, np.float32)
    out = np.zeros(n * n, np.float32)
    sl = mac2d.SemiLagrangian2D()
    with pytest.raises(ValueError):
        sl.advect(n, 1.0, u, v, src, out, -0.1, 0)
    with pytest.raises(ValueError):
        sl.advect(n, 1.0, u, v, src, out, float("nan"), 0)
    with pytest.raises(ValueError):
        sl.advect(n, 1.0, u, v, src, src, 0.1, 0)
    uf = np.full(F, .5, np.float32)
    with pytest.raises(ValueError):
        sl.advect(n, 1.0, uf, v, uf, uf, 0.1, 1)
    with pytest.raises(ValueError):
        sl.advect(n, 1.0, u, v, src, out, 0.1, 7)
    bad = u.copy()
    bad[3] = np.nan
    with pytest.raises(ValueError, match="Non-finite 2D advection input"):
        sl.advect(n, 1.0, bad, v, src, out, 0.1, 0)
    with pytest.raises(ValueError, match="Non-finite 2D backtrace"):
        sl.advect(n, 1.0, u, v, src, out, 1e38, 0)
    sl.advect(n, 1.0, u, v, np.full(n * n, .7, np.float32), out, 100.0, 0)
    assert np.all(out == np.float32(.7))
    say("advection contract enforced")


def test_projection_matches_reference_fixtures_bitwise():
    g = np.load(GOLDEN / "mac2d_stages.npz")
    for k, (n, box, dt, relax, iters, fixed) in enumerate(g["proj_cases"]):
        n, iters = int(n), int(iters)
        o = options(float(dt), float(relax), iters, int(fixed))
        u, v, p, r = gpu_project(n, float(box), o, g[f"proj{k}_u0"], g[f"proj{k}_v0"])
        assert bits_equal(u, g[f"proj{k}_u"]) and bits_equal(v, g[f"proj{k}_v"]), k
        assert bits_equal(p, g[f"proj{k}_p"]), k
        assert bits_equal(row(r), g[f"proj{k}_result"]), (k, row(r), g[f"proj{k}_result"])
        say(f"projection case {k} n={n}: {r.iterations} sweeps, bit-identical (faces, pressure, record)")


@pytest.mark.parametrize("n", [31, 33, 64, 96, 100, 128, 129, 160])
def test_projection_paths_match_oracle_bitwise(n):
    """Tile widths T = 1..4 and the global-memory path (n > 128)."""
    rng = np.random.default_rng(100 + n)
    F = n * (n + 1)
    u = rng.standard_normal(F).astype(np.float32)
    v = rng.standard_normal(F).astype(np.float32)
    for o in (options(1 / 60, 1.9, 600), options(0.05, 1.7, 37, fixed=1), options(1 / 60, 1.95, 200, check=5)):
        gu, gv, gp, gr = gpu_project(n, 4.0, o, u, v)
        rc, ou, ov, op, orr = oracle_project(n, 4.0, o, u, v)
        assert rc == 0
        assert bits_equal(gu, ou) and bits_equal(gv, ov) and bits_equal(gp, op), n
        assert bits_equal(row(gr), row(orr)), (n, row(gr), row(orr))
    say("projection n", n, "bit-identical to the oracle;", gr.iterations, "sweeps in the last case")


def test_projection_failures_leave_the_oracle_state():
    n = 12
    F = n * (n + 1)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(F).astype(np.float32)
    v = rng.standard_normal(F).astype(np.float32)
    bad = u.copy()
    bad[5] = np.inf
    with pytest.raises(ValueError, match="Non-finite 2D velocity"):
        gpu_project(n, 1.0, options(.02, 1.6, 50), bad, v)
    huge = u.copy()
    huge[7 + (n + 1) * 3] = 3.0e38
    huge[8 + (n + 1) * 3] = -3.0e38
    state = mac2d.MacGridState2D(n, 1.0)
    state.velocityX()[:] = huge
    state.velocityY()[:] = v
    with pytest.raises(ValueError, match="Non-finite 2D MAC divergence"):
        mac2d.GpuPressureSolver2D().project(state, options(.02, 1.6, 50))
    rc, ou, ov, _, _ = oracle_project(n, 1.0, options(.02, 1.6, 50), huge, v)
    assert rc == 1 and bits_equal(state.velocityX(), ou) and bits_equal(state.velocityY(), ov)
    for o in (options(0.0, 1.6, 10), options(.02, 2.0, 10), options(.02, 1.6, 10, check=0),
              options(.02, 1.6, -1), options(.02, 1.6, 10, abs_tol=float("nan"))):
        with pytest.raises(ValueError, match="Invalid 2D SOR projection options"):
            gpu_project(n, 1.0, o, u, v)
    say("projection failures: state and error class as the oracle's")


# ----------------------------------------------------------- MacGridFluidSolver2D

def params(**kw):
    p = mac2d.MacGridParameters2D()
    for k, val in kw.items():
        setattr(p, k, val)
    return p


def test_trajectories_match_reference_fixtures():
    g = np.load(GOLDEN / "mac2d_steps.npz")
    for k, (n, box, refl, count) in enumerate(g["runs"]):
        n, count = int(n), int(count)
        sim = mac2d.GpuMacFluidSolver2D(n, float(box), params(reflection=int(refl)))
        orc = O.OracleSim2D(n, float(box), params(reflection=int(refl)))
        for s in range(count):
            sim.advance(DT)
            assert orc.advance(DT) == 0
            for name, got, want in zip(("rho", "u", "v"), sim.fields(), orc.fields()):
                assert bits_equal(got, want), (k, s, name)
                assert close(got, g[f"run{k}_step{s}_{name}"]), (k, s, name)
            assert bits_equal(row(sim.lastProjection()), row(orc.projection())), (k, s)
            assert bits_equal(row(sim.midpointProjection()), row(orc.projection(True))), (k, s)
            assert sim.lastProjection().iterations == int(g[f"run{k}_step{s}_last"][0])
        d = sim.diagnostics()
        rc, od = orc.diagnostics()
        assert rc == 0
        assert bits_equal(np.float32([d.kinetic_energy, d.max_velocity, d.max_divergence]),
                          np.float32([od.kinetic_energy, od.max_velocity, od.max_divergence]))
        say(f"run {k} (n={n}, reflection={int(refl)}): {count} steps bit-identical to the oracle, "
            f"within 2e-5 of vapor's fixtures")


@pytest.mark.parametrize("n,box,refl,steps", [(64, 4.0, 0, 4), (100, 1.0, 1, 3), (128, 1.0, 1, 3), (150, 2.0, 0, 2)])
def test_trajectories_at_larger_grids_match_oracle(n, box, refl, steps):
    p = params(reflection=refl, source_strength=1.0, emitter_radius=0.1, project_iterations=300)
    sim = mac2d.GpuMacFluidSolver2D(n, box, p)
    orc = O.OracleSim2D(n, box, p)
    for s in range(steps):
        sim.advance(DT)
        assert orc.advance(DT) == 0
    for name, got, want in zip(("rho", "u", "v"), sim.fields(), orc.fields()):
        assert bits_equal(got, want), (n, name)
    assert bits_equal(row(sim.lastProjection()), row(orc.projection()))
    say(f"n={n} reflection={refl}: {steps} steps bit-identical; last solve {sim.lastProjection().iterations} sweeps")


def test_default_plume_128_matches_oracle():
    """MacGridFluidSolver2D() (128^2, box 1, plume parameters), 3 steps."""
    sim = mac2d.GpuMacFluidSolver2D()
    p = mac2d.MacGridParameters2D(plume=True)
    orc = O.OracleSim2D(128, 1.0, p)
    for _ in range(3):
        sim.advance(DT)
        assert orc.advance(DT) == 0
    for name, got, want in zip(("rho", "u", "v"), sim.fields(), orc.fields()):
        assert bits_equal(got, want), name
    say("default 128^2 plume: 3 steps bit-identical;", sim.lastProjection().iterations, "sweeps")


def test_step_failures_and_rejections():
    n = 12
    p = params(source_strength=1.0)
    sim = mac2d.GpuMacFluidSolver2D(n, 1.0, p)
    orc = O.OracleSim2D(n, 1.0, p)
    for _ in range(2):
        sim.advance(1 / 30)
        orc.advance(1 / 30)
    for dt in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError):
            sim.advance(dt)
    sim.parameters().sorOmega = 2.0
    with pytest.raises(ValueError):
        sim.advance(DT)
    sim.parameters().sorOmega = 1.9
    sim.parameters().emitterKickMultiplier = -1.0
    with pytest.raises(ValueError, match="kick"):
        sim.advance(DT)
    sim.parameters().emitterKickMultiplier = 1.0
    rho, u, v = sim.fields()
    u[30] = np.nan
    state = mac2d.MacGridState2D(n, 1.0)
    state.density()[:], state.velocityX()[:], state.velocityY()[:] = rho, u, v
    sim.upload(state)
    orc.set_fields(u=u)
    last = row(sim.lastProjection())
    with pytest.raises(ValueError, match="Non-finite 2D advection input"):
        sim.advance(DT)
    assert orc.advance(DT) == 1
    for got, want in zip(sim.fields(), orc.fields()):
        assert bits_equal(got, want)
    assert bits_equal(row(sim.lastProjection()), last)
    say("rejections and a failing step leave the oracle's state")


def test_plume_simulation_mirrors_reference_test():
    """Mac2DTests.cpp:183-227 on the GPU: rest, 120 frames of a 24^2 plume
    with divergence < 0.005, closed walls, bounded density, rising smoke."""
    sim = mac2d.GpuMacFluidSolver2D(24)
    sim.parameters().sourceStrength = 0.0
    for _ in range(5):
        sim.advance(DT)
    assert sim.diagnostics().kinetic_energy == 0.0
    sim.parameters().sourceStrength = 4.0
    n = 24
    for _ in range(120):
        sim.advance(DT)
        assert sim.diagnostics().max_divergence < .005
        rho, u, v = sim.fields()
        assert np.all(u.reshape(n, n + 1)[:, [0, n]] == 0) and np.all(v.reshape(n + 1, n)[[0, n], :] == 0)
        assert np.all(np.isfinite(rho)) and rho.min() >= 0 and rho.max() <= 8.01
    assert sim.diagnostics().kinetic_energy > .001

    def center_y():
        rho = sim.fields()[0].reshape(n, n).astype(np.float64)
        y = (np.arange(n) + .5) * sim.cellSize()
        return float((rho.sum(axis=1) * y).sum() / rho.sum())
    before = center_y()
    sim.parameters().sourceStrength = 0.0
    for _ in range(30):
        sim.advance(DT)
    assert center_y() > before + .02
    sim.resetState()
    assert np.all(sim.fields()[0] == 0) and sim.diagnostics().kinetic_energy == 0.0
    say("plume: divergence, walls, bounds and rise as vapor's own test asserts")


def test_reflection_mirrors_reference_test():
    """Mac2DTests.cpp:229-255: converged midpoint and final projections,
    divergence < 1e-4 on a solenoidal vortex, replay identical, reflection
    differs from simple stepping."""
    n, box = 20, 4.0
    state = mac2d.MacGridState2D(n, box)
    h = state.cellSize()

    def psi(x, y):
        return .8 * np.sin(np.pi * x / 20.) * np.sin(2. * np.pi * y / 20.)
    for y in range(n):
        for x in range(n + 1):
            state.velocityX()[state.idX(x, y)] = np.float32((psi(x, y + 1) - psi(x, y)) / h)
    for y in range(n + 1):
        for x in range(n):
            state.velocityY()[state.idY(x, y)] = np.float32(-(psi(x + 1, y) - psi(x, y)) / h)
    sims = []
    for refl in (1, 0, 1):
        s = mac2d.GpuMacFluidSolver2D.from_state(state)
        s.parameters().sourceStrength = 0.0
        s.parameters().buoyancy = 0.0
        s.parameters().reflection = refl
        sims.append(s)
    reflected, simple, replay = sims
    for _ in range(40):
        for s in sims:
            s.advance(.02)
        assert reflected.midpointProjection().converged and reflected.lastProjection().converged
        assert reflected.diagnostics().max_divergence < 1e-4
    assert bits_equal(reflected.fields()[1], replay.fields()[1])
    assert np.max(np.abs(reflected.fields()[1] - simple.fields()[1])) > 1e-4
    reflected.resetState()
    assert not reflected.midpointProjection().residual_available
    say("reflection: converged, replayable, distinct from simple stepping")


def test_device_fields_and_box_change():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    sim = mac2d.GpuMacFluidSolver2D(16)
    for _ in range(10):
        sim.advance(.02)
    ptrs = sim.deviceFields()
    for name, want in zip(("density", "velocityX", "velocityY"), sim.fields()):
        got = np.empty_like(want)
        assert hip.hipMemcpy(got.ctypes.data, ptrs[name], got.nbytes, 2) == 0
        assert bits_equal(got, want)
    sim.setBoxSize(2.0)
    assert sim.boxSize() == 2.0 and np.all(sim.fields()[0] == 0)
    sim.advance(.02)
    assert sim.lastProjection().residual_available
    say("device addresses hold the state; a box change resets it")


# ----------------------------------------------------- ReflectionMacFluidSolver2D

def hist_params(n, mc, stir):
    p = O.hist_params(n)
    p.maccormack_smoke, p.stir_strength = mc, stir
    return p


def test_historical_matches_reference_fixtures():
    g = np.load(GOLDEN / "mac2d_hist_steps.npz")
    for k, (n, mc, stir, count) in enumerate(g["runs"]):
        n, count = int(n), int(count)
        p = hist_params(n, int(mc), float(stir))
        sim = mac2d.ReflectionMacFluidSolver2D(n, p)
        orc = O.OracleHist2D(n, p)
        for s in range(count):
            sim.advance(DT)
            orc.advance(DT)
            got, want = sim.fields(), orc.fields()
            for name in ("smoke", "vx", "vy", "pressure", "divergence"):
                assert bits_equal(got[name], want[name]), (k, s, name)
            for name in ("smoke", "vx", "vy"):
                assert close(got[name], g[f"run{k}_step{s}_{name}"]), (k, s, name)
            assert sim.randomState() == orc.random_state() == int(g[f"run{k}_step{s}_random"][0])
        say(f"historical run {k} (n={n}, MacCormack={int(mc)}, stir={stir}): bit-identical to the oracle")


@pytest.mark.parametrize("n,mc,stir", [(128, 1, 0.0), (128, 0, 0.4), (150, 1, 0.3), (200, 1, 0.0), (512, 0, 0.0)])
def test_historical_paths_match_oracle(n, mc, stir):
    """LDS and global limiter, tiles T = 4..6 and the global projection."""
    p = hist_params(n, mc, stir)
    p.project_iterations = 40
    sim = mac2d.ReflectionMacFluidSolver2D(n, p)
    orc = O.OracleHist2D(n, p)
    for _ in range(2):
        sim.advance(DT)
        orc.advance(DT)
    got, want = sim.fields(), orc.fields()
    for name in got:
        assert bits_equal(got[name], want[name]), (n, name)
    say(f"historical n={n}: 2 steps bit-identical")


def test_historical_public_steps_mirror_reference_test():
    """Historical2DTests.cpp:24-63: quarter-cell translation, SL vs MC,
    bounds, walls and replay."""
    n = 16
    sl = mac2d.ReflectionMacFluidSolver2D(n, hist_params(n, 0, 0.0))
    mc = mac2d.ReflectionMacFluidSolver2D(n, hist_params(n, 1, 0.0))
    for sim in (sl, mc):
        smoke = np.zeros(n * n, np.float32)
        smoke[7 + n * 8] = 1.0
        sim.upload(smoke=smoke, vx=np.ones(n * (n + 1), np.float32))
        sim.advectScalar(np.float32(.25) * np.float32(1.0 / n))
    s = sl.smoke().reshape(n, n)
    want = np.zeros((n, n), np.float32)
    want[8, 7], want[8, 8] = .75, .25
    assert np.max(np.abs(s - want)) <= 1e-6
    assert not np.array_equal(sl.smoke(), mc.smoke())
    sl.resetState()
    for _ in range(120):
        sl.advance(DT)
        f = sl.fields()
        assert np.all(np.isfinite(f["smoke"])) and f["smoke"].min() >= 0 and f["smoke"].max() <= 1
        assert np.all(f["vx"].reshape(n, n + 1)[:, [0, n]] == 0) and np.all(f["vy"].reshape(n + 1, n)[[0, n], :] == 0)
    first = sl.fields()
    sl.resetState()
    for _ in range(120):
        sl.advance(DT)
    again = sl.fields()
    for name in ("smoke", "vx", "vy"):
        assert bits_equal(first[name], again[name])
    with pytest.raises(ValueError):
        sl.setBoxSize(2.0)
    with pytest.raises(ValueError):
        mac2d.ReflectionMacFluidSolver2D(1)
    with pytest.raises(ValueError):
        sl.advance(0.0)
    say("historical: translation, bounds, walls, replay as vapor's test asserts")


def test_historical_emit_and_project_steps_match_oracle():
    n = 24
    p = hist_params(n, 1, 0.9)
    p.emitter_radius, p.emitter_center_y = 0.3, 0.45
    sim = mac2d.ReflectionMacFluidSolver2D(n, p)
    orc = O.OracleHist2D(n, p)
    rng = np.random.default_rng(8)
    vx = rng.standard_normal(n * (n + 1)).astype(np.float32)
    vy = rng.standard_normal(n * (n + 1)).astype(np.float32)
    sim.upload(vx=vx, vy=vy)
    orc.set_field("vx", vx)
    orc.set_field("vy", vy)
    sim.emit(1.5, DT)
    orc.emit(1.5)
    sim.applyBoundary()
    orc.apply_boundary()
    sim.project(0.02, 25)
    orc.project(0.02, 25)
    got, want = sim.fields(), orc.fields()
    for name in got:
        assert bits_equal(got[name], want[name]), name
    assert sim.randomState() == orc.random_state()
    d, od = sim.diagnostics(), orc.diagnostics()
    assert bits_equal(np.float32([d.kinetic_energy, d.max_velocity, d.max_divergence]),
                      np.float32([od.kinetic_energy, od.max_velocity, od.max_divergence]))
    say("historical emit (stirred), walls, projection and diagnostics bit-identical")
