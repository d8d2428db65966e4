"""CPU build of the device advection code (tests/hostcheck): every field read
is bounds-checked, for the fixtures' inputs and for adversarial ones (NaN,
infinite and huge velocities, n = 1 and 2). The same code must also reproduce
the reference's fixtures bit for bit on the host."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from helpers import bits_equal, load, sizes

LIB = Path(__file__).resolve().parent / "hostcheck" / "_build" / "libhostcheck.so"
_f = ctypes.POINTER(ctypes.c_float)


def lib():
    if not LIB.exists():
        pytest.fail(f"host check library not built: {LIB} (run __graft_entry__.build())")
    h = ctypes.CDLL(str(LIB))
    for fn in (h.hc_advect, h.hc_advect_pow2):
        fn.restype = ctypes.c_long
        fn.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                       _f, _f, _f, _f, _f, _f]
    return h


def p(a):
    return None if a is None else np.ascontiguousarray(a, np.float32).ctypes.data_as(_f)


def advect(n, box, field, method, dt, cfl, flow, src, tsrc=None, pow2=False):
    cells, faces = sizes(n)
    out = np.zeros(cells if field < 0 else faces, np.float32)
    keep = [np.ascontiguousarray(a, np.float32) for a in (*flow, src)] + (
        [np.ascontiguousarray(tsrc, np.float32)] if tsrc is not None else [])
    fn = lib().hc_advect_pow2 if pow2 else lib().hc_advect
    oob = fn(n, box, field, method, dt, cfl, p(keep[0]), p(keep[1]), p(keep[2]), p(keep[3]),
             p(keep[4]) if tsrc is not None else None, out.ctypes.data_as(_f))
    return out, oob


@pytest.mark.parametrize("name", ["stages_n5.npz", "stages_n9.npz"])
def test_device_advection_code_matches_reference_on_host(name):
    g = load(name)
    n, box = int(g["n"]), float(g["box"])
    flow = (g["flow_u"], g["flow_v"], g["flow_w"])
    for di, dt in enumerate(g["dts"]):
        for ci, cfl in enumerate(g["cfls"]):
            for method in (0, 1):
                for axis in (0, 1, 2):
                    out, oob = advect(n, box, axis, method, float(dt), float(cfl), flow, g["src_face"])
                    assert oob == 0
                    assert bits_equal(out, g[f"face_a{axis}_m{method}_d{di}_c{ci}"])
                out, oob = advect(n, box, -1, method, float(dt), float(cfl), flow, g["src_cell"])
                assert oob == 0
                assert bits_equal(out, g[f"cell_m{method}_d{di}_c{ci}"])


@pytest.mark.parametrize("n", [1, 2, 3, 7])
@pytest.mark.parametrize("kind", ["nan", "inf", "huge", "tiny_box"])
def test_device_advection_indices_stay_in_bounds(n, kind):
    rng = np.random.default_rng(n)
    cells, faces = sizes(n)
    flow = [rng.standard_normal(faces).astype(np.float32) * 3 for _ in range(3)]
    box = 1e-30 if kind == "tiny_box" else 1.0
    if kind == "nan":
        for f in flow:
            f[rng.integers(0, faces, max(1, faces // 3))] = np.nan
    elif kind == "inf":
        for f in flow:
            f[rng.integers(0, faces, max(1, faces // 3))] = np.inf
            f[rng.integers(0, faces, max(1, faces // 3))] = -np.inf
    elif kind == "huge":
        for f in flow:
            f *= np.float32(1e30)
    src_f = rng.random(faces).astype(np.float32)
    src_c = rng.random(cells).astype(np.float32)
    for method in (0, 1):
        for dt in (0.0, 0.7, 1e20):
            for axis in (-1, 0, 1, 2):
                src = src_c if axis < 0 else src_f
                _, oob = advect(n, box, axis, method, dt, 3.0, flow, src)
                assert oob == 0, (axis, method, dt)
                if axis >= 0:
                    _, oob = advect(n, box, axis, method, dt, 3.0, flow, src, tsrc=src[::-1].copy())
                    assert oob == 0, (axis, method, dt, "reflected")


@pytest.mark.parametrize("n,box", [(4, 1.0), (8, 1.0), (8, 2.0), (16, 0.5), (2, 1.0)])
@pytest.mark.parametrize("kind", ["plain", "nan", "inf", "huge"])
def test_power_of_two_spacing_path_equals_division(n, box, kind):
    """With h = box / n a power of two the kernels multiply by 1 / h instead of
    dividing by h; every output must be bit-identical to the division path."""
    rng = np.random.default_rng(7 * n)
    cells, faces = sizes(n)
    flow = [rng.standard_normal(faces).astype(np.float32) * 40 for _ in range(3)]
    if kind == "nan":
        for f in flow:
            f[rng.integers(0, faces, max(1, faces // 5))] = np.nan
    elif kind == "inf":
        for f in flow:
            f[rng.integers(0, faces, max(1, faces // 5))] = -np.inf
    elif kind == "huge":
        flow[1] *= np.float32(1e30)
    src_f = rng.standard_normal(faces).astype(np.float32)
    src_c = rng.standard_normal(cells).astype(np.float32)
    for method in (0, 1):
        for dt in (1.0 / 60.0, 0.37, 1e-3):
            for axis in (-1, 0, 1, 2):
                src = src_c if axis < 0 else src_f
                for tsrc in ((None,) if axis < 0 else (None, src[::-1].copy())):
                    a, oa = advect(n, box, axis, method, dt, 3.0, flow, src, tsrc)
                    b, ob = advect(n, box, axis, method, dt, 3.0, flow, src, tsrc, pow2=True)
                    assert oa == 0 and ob == 0
                    assert bits_equal(a, b), (axis, method, dt)


def test_power_of_two_spacing_test():
    """pow2_spacing(h) holds exactly for h = 2^k with 1/h exact and normal."""
    import math
    h = lib()
    h.hc_pow2_spacing.restype = ctypes.c_int
    h.hc_pow2_spacing.argtypes = [ctypes.c_float]
    values = [2.0 ** k for k in range(-149, 128)] + [3.0, 1.0 / 3.0, 1.0 / 96.0, 0.75, 1e-40, 0.0, -0.0, -0.5,
                                                     float("inf"), float("nan"), -float("inf"), 4.5 / 128]
    for v in values:
        f = float(np.float32(v))
        m, e = math.frexp(f) if math.isfinite(f) else (0.0, 0)
        want = f > 0 and math.isfinite(f) and m == 0.5 and -125 < e < 126
        assert h.hc_pow2_spacing(f) == int(want), v


def rb_lib():
    h = lib()
    _d = ctypes.POINTER(ctypes.c_double)
    h.hc_sor_rb.restype = ctypes.c_long
    h.hc_sor_rb.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _f, _f, _f, ctypes.c_int, _d,
                            _d]
    h.hc_rb_coverage.restype = ctypes.c_long
    h.hc_rb_coverage.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    return h


def test_tile_lookup_multipliers_are_exact_for_every_grid():
    """The single-scene sweep finds its tile with multiply-shift reciprocals
    computed on the host; they must equal the integer divisions for every
    block of every grid the library accepts (n^3 < 2^31, so n <= 1290)."""
    h = rb_lib()
    h.hc_rb_geom_check.restype = ctypes.c_long
    h.hc_rb_geom_check.argtypes = [ctypes.c_int]
    assert h.hc_rb_geom_check(1290) == 0


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 9, 63, 64, 65, 70])
def test_fused_sweep_grid_covers_every_cell_once(n):
    counts = np.zeros(n ** 3, np.int32)
    bad = rb_lib().hc_rb_coverage(n, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert bad == 0 and np.all(counts == 1)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 9, 12, 33, 36, 64, 65, 68, 70])
@pytest.mark.parametrize("floor", [6.0, 0.0])
@pytest.mark.parametrize("mode", [0, 1])
def test_fused_sweep_on_host_equals_reference_sweep(n, floor, mode):
    """The fused sweep (mode 1 also forms the FP64 residual of its input), run
    lane by lane with unstaged LDS words holding NaN, equals one reference
    sweep bit for bit."""
    import oracle
    rng = np.random.default_rng(n)
    cells = n ** 3
    b = rng.standard_normal(cells).astype(np.float32)
    x = rng.standard_normal(cells).astype(np.float32)
    want = x.copy()
    oracle.oracle_lib().orc_sor_sweep(n, 1.8, floor, oracle.ptr(b), oracle.ptr(want))
    out = np.full(cells, np.nan, np.float32)
    res = np.zeros(7)
    total = np.zeros(1)
    _d = ctypes.POINTER(ctypes.c_double)
    oob = rb_lib().hc_sor_rb(n, 1.8, floor, 0.0, p(b), p(x), out.ctypes.data_as(_f), mode,
                             res.ctypes.data_as(_d), total.ctypes.data_as(_d))
    assert oob == 0
    assert bits_equal(out, want)
    assert total[0] == pytest.approx(float(np.sum(out, dtype=np.float64)), rel=1e-12, abs=1e-9)
    if mode == 0:
        return
    # mode-1 partials are the FP64 residual of the input
    r = oracle.oracle_lib().orc_residual_l2(n, oracle.ptr(b), oracle.ptr(x))
    got = np.sqrt(max(res[0] - res[5] ** 2 / cells, 0.0))
    assert got == pytest.approx(r, rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("n", [4, 8, 12, 36, 64, 68, 72])
@pytest.mark.parametrize("floor", [6.0, 0.0])
def test_register_path_sweep_on_host_equals_reference_sweep(n, floor):
    """The register path of the fused sweep (right-hand side in registers,
    buffer loads that read 0.0 outside the domain), run lane by lane with
    unstaged LDS words holding NaN and every load bounds-checked, equals one
    reference sweep bit for bit, with a nonzero right-hand-side mean."""
    import oracle
    h = rb_lib()
    _d = ctypes.POINTER(ctypes.c_double)
    h.hc_sor_rb_reg.restype = ctypes.c_long
    h.hc_sor_rb_reg.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _f, _f, _f, _d]
    rng = np.random.default_rng(100 + n)
    cells = n ** 3
    b = rng.standard_normal(cells).astype(np.float32)
    x = rng.standard_normal(cells).astype(np.float32)
    meanf = np.float32(0.375)
    want = x.copy()
    shifted = (b - meanf).astype(np.float32)
    oracle.oracle_lib().orc_sor_sweep(n, 1.8, floor, oracle.ptr(shifted), oracle.ptr(want))
    out = np.full(cells, np.nan, np.float32)
    total = np.zeros(1)
    oob = h.hc_sor_rb_reg(n, 1.8, floor, float(meanf), p(b), p(x), out.ctypes.data_as(_f), total.ctypes.data_as(_d))
    assert oob == 0
    assert bits_equal(out, want)
    assert total[0] == pytest.approx(float(np.sum(out, dtype=np.float64)), rel=1e-12, abs=1e-9)
    # the LDS path on the same input, same bits
    ref = np.full(cells, np.nan, np.float32)
    assert h.hc_sor_rb(n, 1.8, floor, float(meanf), p(b), p(x), ref.ctypes.data_as(_f), 0, None, None) == 0
    assert bits_equal(out, ref)


def test_fast_diagonal_division_is_ieee_for_every_float():
    """The fused sweep divides by the diagonal 6 with an FMA-corrected
    reciprocal (rb_div); it must equal IEEE a / 6 for all 2^32 inputs."""
    h = lib()
    h.hc_rb_div_check.restype = ctypes.c_long
    h.hc_rb_div_check.argtypes = [ctypes.c_float, ctypes.c_uint, ctypes.c_int]
    assert h.hc_rb_div_check(6.0, 1, 8) == 0


@pytest.mark.parametrize("d", [1.0, 2.0, 3.0, 4.0, 5.0])
def test_fast_division_other_diagonals_sampled(d):
    h = lib()
    h.hc_rb_div_check.restype = ctypes.c_long
    h.hc_rb_div_check.argtypes = [ctypes.c_float, ctypes.c_uint, ctypes.c_int]
    assert h.hc_rb_div_check(d, 61, 8) == 0


def test_device_noise_generator_equals_oracle():
    """The device's Philox + Box-Muller code, compiled for the host, gives the
    oracle's variates bit for bit (both use only IEEE +,-,*,/ and sqrt)."""
    import oracle
    h = lib()
    h.hc_noise_normal.restype = None
    h.hc_noise_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _f]
    olib = oracle.oracle_lib()
    for seed, stream in ((0, 0), (12345, 1), (2 ** 40 + 7, 2)):
        count = 20000
        got = np.empty(count, np.float32)
        h.hc_noise_normal(seed, stream, count, got.ctypes.data_as(_f))
        want = np.array([olib.orc_noise_normal(seed, stream, i) for i in range(count)], np.float32)
        assert bits_equal(got, want)
