"""MAC 2D checkers on the CPU.

* The restatement (oracle/mac2d_oracle.c) with the C library's expf/powf
  reproduces vapor's own 2D solvers bit for bit: the committed fixtures
  (tests/golden/mac2d_*.npz, written by gen_golden_mac2d.py from the
  reference build) and, where the reference build is present, fresh runs.
* The IEEE exp/log/pow sequences the GPU and the oracle share stay within one
  unit in the last place of the C library (sampled over the ranges the
  solvers use, and exhaustively over a band around 1), with C99's special
  values.
* The CPU build of the device helpers (tests/hostcheck/mac2d_hostcheck.cpp)
  equals the oracle bit for bit and reads nothing out of range, also for
  adversarial inputs; the limiter's anti-diagonal order equals the
  reference's scan order; the stir's LCG jump-ahead equals sequential draws.
* Closed forms the reference's own tests use (Mac2DTests.cpp): translation of
  a linear field, a manufactured pressure, a solenoidal fixed point, the
  solver's option handling, the historical quarter-cell translation
  (Historical2DTests.cpp:24-38).
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.filterwarnings("ignore:overflow encountered in cast")

from oracle import mac2d as O
from vapor_amd.abi import VgxProjectionOptions, VgxProjectionResult, Vgx2Params, Vgx2rParams, load_library

GOLDEN = Path(__file__).resolve().parent / "golden"
HOSTCHECK = Path(__file__).resolve().parent / "hostcheck" / "_build" / "libhostcheck2d.so"
_f = ctypes.POINTER(ctypes.c_float)
RESULT_FIELDS = ("iterations", "converged", "initial_residual", "final_residual", "residual_available",
                 "divergence_before", "divergence_after", "removed_rhs_mean")


def fixture(name):
    return np.load(GOLDEN / name)


def row(r):
    return np.array([getattr(r, f) for f in RESULT_FIELDS], np.float64)


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, a.dtype)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def params2d(**kw):
    p = Vgx2Params()
    load_library().vgx2_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def hist_params(n, mc, stir):
    p = O.hist_params(n)
    p.maccormack_smoke, p.stir_strength = mc, stir
    return p


def options(dt, relaxation, iters, fixed=0, check=8, abs_tol=1e-6, rel_tol=1e-5):
    o = VgxProjectionOptions()
    o.dt, o.relaxation, o.max_iterations, o.residual_check_interval = dt, relaxation, iters, check
    o.absolute_tolerance, o.relative_tolerance, o.fixed_iterations = abs_tol, rel_tol, fixed
    return o


# ----------------------------------------------------------- fixtures (libm mode)

def test_oracle_stages_match_reference_fixtures():
    g = fixture("mac2d_stages.npz")
    lib = O.oracle2d_lib()
    n, box = 9, 1.0
    u, v = g["adv_u"], g["adv_v"]
    with O.libm_mode():
        for field in (0, 1, 2):
            src = g[f"adv_src{field}"]
            for k, dt in enumerate(g["adv_dts"]):
                out = np.empty_like(src)
                assert lib.o2_advect(n, box, field, float(dt), O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(out)) == 0
                assert bits_equal(out, g[f"adv_out{field}_{k}"]), (field, dt)
        for k, (n, box, dt, relax, iters, fixed) in enumerate(g["proj_cases"]):
            n, iters = int(n), int(iters)
            uu, vv = g[f"proj{k}_u0"].copy(), g[f"proj{k}_v0"].copy()
            p = np.zeros(n * n, np.float64)
            r = VgxProjectionResult()
            o = options(float(dt), float(relax), iters, int(fixed))
            assert lib.o2_project(n, float(box), ctypes.byref(o), O.fptr(uu), O.fptr(vv), O.dptr(p),
                                  ctypes.byref(r)) == 0
            assert bits_equal(uu, g[f"proj{k}_u"]) and bits_equal(vv, g[f"proj{k}_v"]), k
            assert bits_equal(p, g[f"proj{k}_p"]), k
            assert bits_equal(row(r), g[f"proj{k}_result"]), (k, row(r), g[f"proj{k}_result"])


def test_oracle_trajectories_match_reference_fixtures():
    g = fixture("mac2d_steps.npz")
    with O.libm_mode():
        for k, (n, box, refl, count) in enumerate(g["runs"]):
            n, count = int(n), int(count)
            sim = O.OracleSim2D(n, float(box), params2d(reflection=int(refl)))
            for s in range(count):
                assert sim.advance(1 / 60) == 0
                for name, got in zip(("rho", "u", "v"), sim.fields()):
                    assert bits_equal(got, g[f"run{k}_step{s}_{name}"]), (k, s, name)
                assert bits_equal(row(sim.projection()), g[f"run{k}_step{s}_last"]), (k, s)
                assert bits_equal(row(sim.projection(True)), g[f"run{k}_step{s}_mid"]), (k, s)
            rc, d = sim.diagnostics()
            want = g[f"run{k}_diag"]
            assert rc == want[0]
            assert bits_equal(np.array([d.kinetic_energy, d.max_velocity, d.max_divergence], np.float32),
                              want[1:].astype(np.float32))


def test_oracle_historical_trajectories_match_reference_fixtures():
    g = fixture("mac2d_hist_steps.npz")
    with O.libm_mode():
        for k, (n, mc, stir, count) in enumerate(g["runs"]):
            n, count = int(n), int(count)
            sim = O.OracleHist2D(n, hist_params(n, int(mc), float(stir)))
            for s in range(count):
                assert sim.advance(1 / 60) == 0
                f = sim.fields()
                for name in ("smoke", "vx", "vy", "pressure"):
                    assert bits_equal(f[name], g[f"run{k}_step{s}_{name}"]), (k, s, name)
                assert sim.random_state() == int(g[f"run{k}_step{s}_random"][0])
            d = sim.diagnostics()
            assert bits_equal(np.array([d.kinetic_energy, d.max_velocity, d.max_divergence], np.float32),
                              g[f"run{k}_diag"].astype(np.float32))


@pytest.mark.skipif(not O.ref2d_available(), reason="reference 2D build absent (no reference tree here)")
def test_oracle_matches_fresh_reference_runs_with_errors():
    """Cases the fixtures leave out: rejections and failures inside a step
    leave both solvers in the same state with the same error class."""
    with O.libm_mode():
        for refl in (0, 1):
            prm = params2d(reflection=refl, source_strength=1.0)
            ref, orc = O.RefSim2D(12, 1.0, prm), O.OracleSim2D(12, 1.0, prm)
            for _ in range(3):
                assert ref.advance(1 / 30) == orc.advance(1 / 30) == 0
            for dt in (0.0, -1.0, float("nan")):
                assert ref.advance(dt) == orc.advance(dt) == 1
            # a NaN face makes the velocity advection throw before any change
            rho, u, v = orc.fields()
            u[40] = np.nan
            orc.set_fields(u=u)
            ref2 = O.RefSim2D(12, 1.0, prm, rho, u, v)
            assert ref2.advance(1 / 30) == orc.advance(1 / 30) == 1
            for a, b in zip(ref2.fields(), orc.fields()):
                assert bits_equal(a, b)


# ------------------------------------------------------------- transcendentals

def ulp_distance(a, b):
    ia = np.frombuffer(np.asarray(a, np.float32).tobytes(), np.int32).astype(np.int64)
    ib = np.frombuffer(np.asarray(b, np.float32).tobytes(), np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def sweep_floats(lo, hi, step):
    a = np.frombuffer(np.float32(lo).tobytes(), np.uint32)[0]
    b = np.frombuffer(np.float32(hi).tobytes(), np.uint32)[0]
    lo_i, hi_i = min(a, b), max(a, b)
    return np.frombuffer(np.arange(lo_i, hi_i + 1, step, dtype=np.uint32).tobytes(), np.float32)


def test_shared_exp_within_one_ulp_of_libm():
    lib = O.oracle2d_lib()
    xs = np.concatenate([-sweep_floats(0.0, 3.5, 997), sweep_floats(-88.0, -3.5, 4999) * 1,
                         sweep_floats(0.0, 88.0, 4999)])
    worst = 0
    with O.libm_mode():
        libm = np.array([lib.o2_expf(float(x)) for x in xs], np.float32)
    det = np.array([lib.o2_expf(float(x)) for x in xs], np.float32)
    worst = int(ulp_distance(libm, det).max())
    assert worst <= 1, worst
    # exact at the points where e^x is known
    assert lib.o2_expf(0.0) == 1.0 and lib.o2_expf(float("-inf")) == 0.0 and lib.o2_expf(float("inf")) == float("inf")
    assert np.isnan(lib.o2_expf(float("nan")))
    # agreement with a double-precision exp on a fine grid of [-3.5, 0]: the
    # device value is the correctly rounded one except within 1e-12 of a tie
    grid = -sweep_floats(0.0, 3.5, 131)
    det = np.array([lib.o2_expf(float(x)) for x in grid], np.float32)
    assert np.mean(det == np.exp(grid.astype(np.float64)).astype(np.float32)) > 0.99999


def test_shared_pow_within_one_ulp_of_libm_and_special_values():
    lib = O.oracle2d_lib()
    xs = np.concatenate([sweep_floats(0.0, 1.0, 1009), sweep_floats(1.0, 64.0, 2003)])
    for y in (0.25, 0.5, 1.5, -0.25, 3.0, 2.7):
        with O.libm_mode():
            libm = np.array([lib.o2_powf(float(x), y) for x in xs], np.float32)
        det = np.array([lib.o2_powf(float(x), y) for x in xs], np.float32)
        finite = np.isfinite(libm)
        assert np.array_equal(np.isfinite(det), finite)
        assert int(ulp_distance(libm[finite], det[finite]).max()) <= 1, y
    inf, nan = float("inf"), float("nan")
    special = [(0.0, 0.25), (-0.0, 0.25), (-0.0, 3.0), (0.0, -1.0), (-0.0, -3.0), (1.0, nan), (nan, 0.0),
               (-1.0, inf), (0.5, inf), (2.0, inf), (0.5, -inf), (inf, 0.25), (-inf, 3.0), (-inf, 2.0),
               (-inf, -3.0), (-8.0, 1.0 / 3.0), (-2.0, 3.0), (-2.0, 2.0), (nan, 1.0), (2.0, nan)]
    for x, y in special:
        with O.libm_mode():
            want = lib.o2_powf(x, y)
        got = lib.o2_powf(x, y)
        if np.isnan(want):
            assert np.isnan(got), (x, y)
        else:
            assert bits_equal(np.float32(got), np.float32(want)), (x, y, got, want)


def test_quarter_power_exhaustive_band():
    """x^(1/4) for every float in [0.0625, 1]: the buoyancy's range."""
    lib = O.oracle2d_lib()
    xs = sweep_floats(0.0625, 1.0, 61)
    with O.libm_mode():
        libm = np.array([lib.o2_powf(float(x), 0.25) for x in xs], np.float32)
    det = np.array([lib.o2_powf(float(x), 0.25) for x in xs], np.float32)
    assert int(ulp_distance(libm, det).max()) <= 1
    # four exact quarter powers
    for x, r in ((16.0, 2.0), (0.0625, 0.5), (1.0, 1.0), (81.0, 3.0)):
        assert lib.o2_powf(x, 0.25) == r


def test_device_mode_trajectory_stays_close_to_reference():
    """The transcendental substitution moves trajectories by float rounding only."""
    g = fixture("mac2d_steps.npz")
    for k, (n, box, refl, count) in enumerate(g["runs"]):
        sim = O.OracleSim2D(int(n), float(box), params2d(reflection=int(refl)))
        for s in range(int(count)):
            assert sim.advance(1 / 60) == 0
        for name, got in zip(("rho", "u", "v"), sim.fields()):
            want = g[f"run{k}_step{int(count) - 1}_{name}"]
            assert np.max(np.abs(got.astype(np.float64) - want)) <= 2e-5 * max(1.0, float(np.max(np.abs(want))))
    g = fixture("mac2d_hist_steps.npz")
    for k, (n, mc, stir, count) in enumerate(g["runs"]):
        sim = O.OracleHist2D(int(n), hist_params(int(n), int(mc), float(stir)))
        for s in range(int(count)):
            sim.advance(1 / 60)
        f = sim.fields()
        for name in ("smoke", "vx", "vy"):
            want = g[f"run{k}_step{int(count) - 1}_{name}"]
            assert np.max(np.abs(f[name].astype(np.float64) - want)) <= 2e-5 * max(1.0, float(np.max(np.abs(want))))


# ---------------------------------------------------------- device code on host

def hostcheck():
    if not HOSTCHECK.exists():
        pytest.fail(f"host check library not built: {HOSTCHECK} (run __graft_entry__.build())")
    h = ctypes.CDLL(str(HOSTCHECK))
    h.hc2_det_exp.restype = h.hc2_det_log.restype = ctypes.c_double
    h.hc2_det_exp.argtypes = h.hc2_det_log.argtypes = [ctypes.c_double]
    h.hc2_expf.restype = ctypes.c_float
    h.hc2_expf.argtypes = [ctypes.c_float]
    h.hc2_powf.restype = h.hc2_lerp.restype = ctypes.c_float
    h.hc2_powf.argtypes = [ctypes.c_float, ctypes.c_float]
    h.hc2_lerp.argtypes = [ctypes.c_float] * 3
    h.hc2_lcg_jump.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    h.hc2_advect.restype = h.hc2_hist_trace.restype = h.hc2_hist_face.restype = h.hc2_limit.restype = ctypes.c_long
    h.hc2_advect.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float, _f, _f, _f, _f]
    h.hc2_hist_trace.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int, _f, _f, _f, _f]
    h.hc2_hist_face.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int, _f, _f, _f, _f]
    h.hc2_limit.argtypes = [ctypes.c_int, _f, _f, _f]
    h.hc2_hist_stir.restype = ctypes.c_uint32
    h.hc2_hist_stir.argtypes = [ctypes.c_int] + [ctypes.c_float] * 5 + [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, _f]
    return h


def test_device_transcendentals_equal_oracle_bitwise():
    h, lib = hostcheck(), O.oracle2d_lib()
    rng = np.random.default_rng(5)
    for t in np.concatenate([rng.uniform(-745, 709, 2000), rng.uniform(-4, 1, 2000), [0.0, -0.0, 1e-300]]):
        assert h.hc2_det_exp(float(t)) == lib.o2_det_exp(float(t))
    for x in np.concatenate([np.exp(rng.uniform(-700, 700, 2000)), [1.0, 2.0, 0.5, 3.0e-300]]):
        assert h.hc2_det_log(float(x)) == lib.o2_det_log(float(x))
    xs = np.concatenate([rng.random(3000).astype(np.float32), (rng.random(300) * 50).astype(np.float32),
                         np.float32([0, -0.0, 1, np.inf, -np.inf, np.nan, -2, -8])])
    for y in (0.25, 0.5, -1.0, 3.0, np.inf, -np.inf, 0.0, np.nan):
        for x in xs:
            a, b = np.float32(h.hc2_powf(float(x), float(y))), np.float32(lib.o2_powf(float(x), float(y)))
            assert (np.isnan(a) and np.isnan(b)) or bits_equal(a, b), (x, y)
    for v in rng.uniform(-90, 90, 3000).astype(np.float32):
        assert bits_equal(np.float32(h.hc2_expf(float(v))), np.float32(lib.o2_expf(float(v))))


def test_division_by_three_without_a_divide():
    """The projections' FMA-corrected a/3 equals IEEE division (every 61st
    float here; all 2^32 floats and 5e8 doubles were checked once)."""
    h = hostcheck()
    h.hc2_div3_check.restype = ctypes.c_long
    h.hc2_div3_check.argtypes = [ctypes.c_uint32, ctypes.c_long]
    assert h.hc2_div3_check(61, 2_000_000) == 0


def test_device_lerp_is_libstdcxx_lerp():
    """Cases of each branch, against numpy restating P0811's three forms."""
    h = hostcheck()
    rng = np.random.default_rng(9)
    for a, b, t in zip(rng.standard_normal(4000).astype(np.float32), rng.standard_normal(4000).astype(np.float32),
                       rng.random(4000).astype(np.float32)):
        got = np.float32(h.hc2_lerp(float(a), float(b), float(t)))
        if (a <= 0 and b >= 0) or (a >= 0 and b <= 0):
            want = np.float32(np.float32(t * b) + np.float32(np.float32(1 - t) * a))
        else:
            x = np.float32(a + np.float32(t * np.float32(b - a)))
            want = (x if b < x else b) if ((t > 1) == (b > a)) else (x if b > x else b)
        assert bits_equal(got, want), (a, b, t)
    assert h.hc2_lerp(2.0, 5.0, 1.0) == 5.0 and h.hc2_lerp(2.0, 5.0, 0.0) == 2.0


def test_device_samplers_equal_oracle_and_stay_in_bounds():
    h, lib = hostcheck(), O.oracle2d_lib()
    g = fixture("mac2d_stages.npz")
    n, box = 9, 1.0
    u, v = g["adv_u"], g["adv_v"]
    with O.libm_mode():
        for field in (0, 1, 2):
            src = g[f"adv_src{field}"]
            for k, dt in enumerate(g["adv_dts"]):
                out = np.empty_like(src)
                assert h.hc2_advect(n, box, field, float(dt), O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(out)) == 0
                assert bits_equal(out, g[f"adv_out{field}_{k}"])
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 17):
        N, F = n * n, n * (n + 1)
        for scale in (1.0, 1e6, 3e38):
            uu = (rng.standard_normal(F) * scale).astype(np.float32)
            vv = (rng.standard_normal(F) * scale).astype(np.float32)
            for field, count in ((0, N), (1, F), (2, F)):
                src = rng.random(count).astype(np.float32)
                got, want = np.empty_like(src), np.empty_like(src)
                rc = h.hc2_advect(n, 1.0, field, 0.5, O.fptr(uu), O.fptr(vv), O.fptr(src), O.fptr(got))
                orc = lib.o2_advect(n, 1.0, field, 0.5, O.fptr(uu), O.fptr(vv), O.fptr(src), O.fptr(want))
                assert rc in (0, -1)
                assert (rc == -1) == (orc == 1), (n, scale, field)
                if rc == 0:
                    assert bits_equal(got, want)


def test_historical_device_samplers_and_limiter_equal_oracle_scan():
    h = hostcheck()
    rng = np.random.default_rng(11)
    for n in (2, 3, 16, 25):
        N, F = n * n, n * (n + 1)
        for scale in (0.5, 40.0, np.float32(3e38)):
            vx = (rng.standard_normal(F) * scale).astype(np.float32)
            vy = (rng.standard_normal(F) * scale).astype(np.float32)
            smoke = rng.random(N).astype(np.float32)
            nan_vx = vx.copy()
            nan_vx[rng.integers(0, F)] = np.nan
            for flow_x in (vx, nan_vx):
                fwd = np.empty(N, np.float32)
                back = np.empty(N, np.float32)
                assert h.hc2_hist_trace(n, 0.02, -1, O.fptr(smoke), O.fptr(flow_x), O.fptr(vy), O.fptr(fwd)) == 0
                assert h.hc2_hist_trace(n, 0.02, 1, O.fptr(fwd), O.fptr(flow_x), O.fptr(vy), O.fptr(back)) == 0
                for axis in (0, 1):
                    out = np.empty(F, np.float32)
                    src = vx if axis == 0 else vy
                    assert h.hc2_hist_face(n, 0.02, axis, O.fptr(src), O.fptr(flow_x), O.fptr(vy), O.fptr(out)) == 0
            if not np.all(np.isfinite(vx)) or scale > 1e30:
                continue  # the reference (and oracle) index with int(NaN) there
            # the limiter: anti-diagonal order against the reference's scan
            oracle = O.OracleHist2D(n, hist_params(n, 1, 0.0))
            oracle.set_field("smoke", smoke)
            oracle.set_field("vx", vx)
            oracle.set_field("vy", vy)
            with O.libm_mode():
                oracle.advect_smoke(0.02)
            want = oracle.fields()["smoke"]
            fwd = np.empty(N, np.float32)
            back = np.empty(N, np.float32)
            h.hc2_hist_trace(n, 0.02, -1, O.fptr(smoke), O.fptr(vx), O.fptr(vy), O.fptr(fwd))
            h.hc2_hist_trace(n, 0.02, 1, O.fptr(fwd), O.fptr(vx), O.fptr(vy), O.fptr(back))
            corr = (fwd + np.float32(0.5) * (smoke - back)).astype(np.float32)
            field = smoke.copy()
            assert h.hc2_limit(n, O.fptr(field), O.fptr(fwd), O.fptr(corr)) == 0
            assert bits_equal(field, want), (n, scale)


def test_stir_jump_ahead_equals_sequential_draws():
    h = hostcheck()
    n = 24
    for stir, cy, r in ((0.7, 0.2, 0.08), (1.3, 0.5, 0.3), (0.2, 0.95, 0.1)):
        p = hist_params(n, 1, stir)
        p.emitter_center_y, p.emitter_radius = cy, r
        orc = O.OracleHist2D(n, p)
        rng = np.random.default_rng(1)
        vx0 = rng.standard_normal(n * (n + 1)).astype(np.float32)
        orc.set_field("vx", vx0)
        orc.set_random_state(0x12345678)
        orc.emit(1.0)
        want = orc.fields()["vx"]
        fn = np.float32(n)
        y_end = min(int(np.float32(np.float32(cy) * fn + np.float32(r) * fn) + np.float32(2)), n)
        y_begin = max(0, int(np.float32(np.float32(cy) * fn - np.float32(r) * fn) - np.float32(1)))
        got = vx0.copy()
        state = h.hc2_hist_stir(n, 0.5, cy, r, 3.5, stir, y_begin, y_end, 0x12345678, O.fptr(got))
        assert bits_equal(got, want)
        assert state == orc.random_state()
    A, C = ctypes.c_uint32(), ctypes.c_uint32()
    s = 0x9E3779B9
    seq = s
    for k in range(1, 300):
        seq = (seq * 1664525 + 1013904223) & 0xFFFFFFFF
        h.hc2_lcg_jump(k, ctypes.byref(A), ctypes.byref(C))
        assert (A.value * s + C.value) & 0xFFFFFFFF == seq


# ---------------------------------------------------------------- closed forms

def test_translation_of_a_linear_field():
    """Mac2DTests.cpp:66-82: exact bilinear translation in world units."""
    lib = O.oracle2d_lib()
    n, h = 8, 1.0 / 8
    F = n * (n + 1)
    u = np.full(F, 0.5, np.float32)
    v = np.full(F, -0.25, np.float32)
    dt = 0.125
    for field in (0, 1, 2):
        w, hgt = n + (field == 1), n + (field == 2)
        ox = 0.0 if field == 1 else 0.5
        oy = 0.0 if field == 2 else 0.5
        x, y = np.meshgrid(np.arange(w), np.arange(hgt))
        src = (1 + 0.2 * (x + ox) * h + 0.3 * (y + oy) * h).astype(np.float32).ravel()
        out = np.empty_like(src)
        assert lib.o2_advect(n, 1.0, field, 0.0, O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(out)) == 0
        assert np.max(np.abs(out - src)) <= 1e-6
        assert lib.o2_advect(n, 1.0, field, dt, O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(out)) == 0
        expect = 1 + 0.2 * ((x + ox) * h - 0.5 * dt) + 0.3 * ((y + oy) * h + 0.25 * dt)
        inner = (slice(1, hgt - 1), slice(1, w - 1))
        assert np.max(np.abs(out.reshape(hgt, w)[inner] - expect[inner])) <= 3e-7
        assert out.min() >= src.min() and out.max() <= src.max()
        assert lib.o2_advect(n, 1.0, field, -dt, O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(out)) == 1
        assert lib.o2_advect(n, 1.0, field, dt, O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(src)) == 1


@pytest.mark.parametrize("n", [1, 8, 15])
@pytest.mark.parametrize("box", [1.0, 3.0])
@pytest.mark.parametrize("dt", [0.01, 0.1])
def test_manufactured_pressure_is_recovered(n, box, dt):
    """Mac2DTests.cpp:93-121: a pure gradient is removed and p recovers phi."""
    lib = O.oracle2d_lib()
    x = (np.arange(n, dtype=np.float32) + np.float32(.5)) / np.float32(n)
    exact = (np.cos(np.float32(np.pi) * x)[None, :] * np.cos(np.float32(2 * np.pi) * x)[:, None]).astype(np.float32)
    exact = (exact - np.float32(exact.astype(np.float64).mean())).ravel()
    h = np.float32(box) / np.float32(n)
    u = np.zeros(n * (n + 1), np.float32)
    v = np.zeros(n * (n + 1), np.float32)
    e = exact.reshape(n, n)
    u.reshape(n, n + 1)[:, 1:n] = np.float32(dt) / h * (e[:, 1:] - e[:, :-1])
    v.reshape(n + 1, n)[1:n, :] = np.float32(dt) / h * (e[1:, :] - e[:-1, :])
    p = np.zeros(n * n, np.float64)
    r = VgxProjectionResult()
    o = options(dt, 1.6, 4000, abs_tol=3e-6, rel_tol=1e-6)
    assert lib.o2_project(n, box, ctypes.byref(o), O.fptr(u), O.fptr(v), O.dptr(p), ctypes.byref(r)) == 0
    assert r.residual_available and r.converged
    assert np.max(np.abs(p - exact)) <= 8e-5
    assert r.divergence_after < 2e-4
    assert np.max(np.abs(u)) <= 3e-5 and np.max(np.abs(v)) <= 3e-5


def test_solenoidal_field_is_a_fixed_point():
    """Mac2DTests.cpp:122-139."""
    lib = O.oracle2d_lib()
    n, box = 12, 2.0
    h = np.float32(box / n)

    def psi(x, y):
        if x == 0 or y == 0 or x == n or y == n:
            return np.float32(0)
        return np.float32(np.sin(np.float32(np.pi) * x / n) * np.sin(np.float32(np.pi) * y / n))
    u = np.zeros(n * (n + 1), np.float32)
    v = np.zeros(n * (n + 1), np.float32)
    for y in range(n):
        for x in range(n + 1):
            u[x + (n + 1) * y] = (psi(x, y + 1) - psi(x, y)) / h
    for y in range(n + 1):
        for x in range(n):
            v[x + n * y] = -(psi(x + 1, y) - psi(x, y)) / h
    u0, v0 = u.copy(), v.copy()
    r = VgxProjectionResult()
    o = options(0.02, 1.6, 4000, abs_tol=3e-6, rel_tol=1e-6)
    assert lib.o2_project(n, box, ctypes.byref(o), O.fptr(u), O.fptr(v), None, ctypes.byref(r)) == 0
    assert np.max(np.abs(u - u0)) <= 3e-6 and np.max(np.abs(v - v0)) <= 3e-6


def test_projection_options_and_budgets():
    """Mac2DTests.cpp:140-180 (rest exits early; fixed budget; zero budget)."""
    lib = O.oracle2d_lib()
    n = 8
    F = n * (n + 1)
    u = np.zeros(F, np.float32)
    v = np.zeros(F, np.float32)
    r = VgxProjectionResult()
    assert lib.o2_project(n, 4.0, ctypes.byref(options(0.02, 1.6, 4000, abs_tol=3e-6, rel_tol=1e-6)), O.fptr(u),
                          O.fptr(v), None, ctypes.byref(r)) == 0
    assert r.iterations == 0
    assert lib.o2_project(n, 4.0, ctypes.byref(options(0.02, 1.6, 17, fixed=1)), O.fptr(u), O.fptr(v), None,
                          ctypes.byref(r)) == 0
    assert r.iterations == 17
    u[4 + (n + 1) * 3] = 1.0
    assert lib.o2_project(n, 4.0, ctypes.byref(options(0.02, 1.6, 0)), O.fptr(u), O.fptr(v), None,
                          ctypes.byref(r)) == 0
    assert not r.converged
    for bad in (options(0.0, 1.6, 10), options(-1.0, 1.6, 10), options(float("inf"), 1.6, 10),
                options(0.02, 2.0, 10), options(0.02, 1.6, 10, check=0), options(0.02, 1.6, -1),
                options(0.02, 1.6, 10, abs_tol=float("nan"))):
        assert lib.o2_project(n, 4.0, ctypes.byref(bad), O.fptr(u), O.fptr(v), None, ctypes.byref(r)) == 1


def test_historical_quarter_cell_translation():
    """Historical2DTests.cpp:24-38: semi-Lagrangian moves a spike a quarter cell."""
    n = 16
    sl = O.OracleHist2D(n, hist_params(n, 0, 0.0))
    mc = O.OracleHist2D(n, hist_params(n, 1, 0.0))
    for sim in (sl, mc):
        sim.set_field("vx", np.ones(n * (n + 1), np.float32))
        smoke = np.zeros(n * n, np.float32)
        smoke[7 + n * 8] = 1.0
        sim.set_field("smoke", smoke)
        sim.advect_smoke(np.float32(0.25) * np.float32(1.0 / n))
    got = sl.fields()["smoke"].reshape(n, n)
    want = np.zeros((n, n), np.float32)
    want[8, 7], want[8, 8] = 0.75, 0.25
    assert np.max(np.abs(got - want)) <= 1e-6
    assert not np.array_equal(sl.fields()["smoke"], mc.fields()["smoke"])
