"""The oracle against what the mathematics fixes, following the reference's own
closed-form and invariant tests (AdvectionTests.cpp, NumericsTests.cpp)."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import f32, max_abs, options, oracle_face, oracle_project, oracle_scalar, ptr, sizes


def linear_fields(n, box=None):
    """x + 2y + 3z in cell units, sampled at each field's points
    (AdvectionTests.cpp:29-40)."""
    value = lambda x, y, z: x + 2 * y + 3 * z  # noqa: E731
    cells, faces = sizes(n)
    rho = np.zeros(cells, np.float32)
    u, v, w = (np.zeros(faces, np.float32) for _ in range(3))
    for z in range(n):
        for y in range(n):
            for x in range(n):
                rho[x + n * (y + n * z)] = value(x + .5, y + .5, z + .5)
    for z in range(n):
        for y in range(n):
            for x in range(n + 1):
                u[x + (n + 1) * (y + n * z)] = value(x, y + .5, z + .5)
    for z in range(n):
        for y in range(n + 1):
            for x in range(n):
                v[x + n * (y + (n + 1) * z)] = value(x + .5, y, z + .5)
    for z in range(n + 1):
        for y in range(n):
            for x in range(n):
                w[x + n * (y + n * z)] = value(x + .5, y + .5, z)
    return rho, u, v, w


def test_sampling_exact_on_linear_fields():
    # AdvectionTests.cpp:29-56: 4^3 grid on box 4 (h = 1)
    lib = oracle.oracle_lib()
    rho, u, v, w = linear_fields(4)
    p = (1.2, 1.6, 2.1)
    expected = p[0] + 2 * p[1] + 3 * p[2]
    for axis, f in ((-1, rho), (0, u), (1, v), (2, w)):
        assert abs(lib.orc_sample(4, 4.0, axis, ptr(f), *p) - expected) <= 1e-5
    assert abs(lib.orc_sample(4, 4.0, -1, ptr(rho), -1, -1, -1) - 3.0) <= 1e-5
    assert abs(lib.orc_sample(4, 4.0, 0, ptr(u), -1, -1, -1) - 2.5) <= 1e-5
    assert abs(lib.orc_sample(4, 4.0, -1, ptr(rho), 9, 9, 9) - 21.0) <= 1e-5
    assert abs(lib.orc_sample(4, 4.0, 0, ptr(u), 9, 9, 9) - 21.5) <= 1e-5
    assert abs(lib.orc_sample(4, 4.0, 1, ptr(v), 9, 9, 9) - 22.0) <= 1e-5
    assert abs(lib.orc_sample(4, 4.0, 2, ptr(w), 9, 9, 9) - 22.5) <= 1e-5
    single = np.array([2.0], np.float32)
    assert lib.orc_sample(1, 4.5, -1, ptr(single), 1, 1, 1) == 2.0


@pytest.mark.parametrize("n", [1, 4, 5])
@pytest.mark.parametrize("method", [0, 1])
def test_zero_velocity_is_identity(n, method):
    # AdvectionTests.cpp:58-82
    cells, faces = sizes(n)
    flow = [np.zeros(faces, np.float32)] * 3
    rho = (np.arange(cells) % 7 / 7).astype(np.float32)
    assert max_abs(oracle_scalar(n, float(n), method, 0.2, 3.0, flow, rho), rho) <= 1e-5
    face = (np.arange(faces) % 13 / 13).astype(np.float32)
    for axis in range(3):
        assert max_abs(oracle_face(n, float(n), axis, method, 0.2, 3.0, flow, face), face) <= 1e-5


@pytest.mark.parametrize("method", [0, 1])
def test_uniform_translation(method):
    # AdvectionTests.cpp:84-110: u = 0.5, dt = 0.5 on 6^3, box 6 -> 12 - 0.25 at (2,2,2)
    n = 6
    cells, faces = sizes(n)
    flow = [np.full(faces, 0.5, np.float32), np.zeros(faces, np.float32), np.zeros(faces, np.float32)]
    rho = np.array([x + 2 * y + 3 * z for z in range(n) for y in range(n) for x in range(n)], np.float32)
    out = oracle_scalar(n, 6.0, method, 0.5, 3.0, flow, rho)
    assert abs(out[2 + 6 * (2 + 6 * 2)] - 11.75) <= 1e-5
    for axis in range(3):
        ex, ey, ez = n + (axis == 0), n + (axis == 1), n + (axis == 2)
        src = np.zeros(faces, np.float32)
        for z in range(ez):
            for y in range(ey):
                for x in range(ex):
                    src[x + ex * (y + ey * z)] = x + 2 * y + 3 * z
        out = oracle_face(n, 6.0, axis, method, 0.5, 3.0, flow, src)
        assert abs(out[2 + ex * (2 + ey * 2)] - 11.75) <= 1e-5


def test_limiter_bounds_symmetry_and_fallback():
    # AdvectionTests.cpp:112-157
    n = 10
    cells, faces = sizes(n)
    right = [np.full(faces, .7, np.float32), np.zeros(faces, np.float32), np.zeros(faces, np.float32)]
    left = [np.full(faces, -.7, np.float32), np.zeros(faces, np.float32), np.zeros(faces, np.float32)]
    rr = np.zeros(cells, np.float32)
    ll = np.zeros(cells, np.float32)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                val = ((x * 17 + y * 7 + z * 3) % 13) / 12
                rr[x + n * (y + n * z)] = val
                ll[(9 - x) + n * (y + n * z)] = val
    fallback = oracle_scalar(n, 10.0, 1, .6, .001, right, rr)
    sl = oracle_scalar(n, 10.0, 0, .6, 3.0, right, rr)
    assert max_abs(fallback, sl) <= 1e-5
    mr = oracle_scalar(n, 10.0, 1, .6, 3.0, right, rr).reshape(n, n, n)
    ml = oracle_scalar(n, 10.0, 1, .6, 3.0, left, ll).reshape(n, n, n)
    assert max_abs(mr[2:8, 2:8, 2:8], ml[2:8, 2:8, ::-1][:, :, 2:8]) <= 2e-6
    n = 6
    cells, faces = sizes(n)
    flow = [np.full(faces, .5, np.float32), np.zeros(faces, np.float32), np.zeros(faces, np.float32)]
    rho = np.zeros(cells, np.float32)
    rho[3 + 6 * (3 + 6 * 3)] = 1
    semi = oracle_scalar(n, 6.0, 0, .4, 3.0, flow, rho)
    mc = oracle_scalar(n, 6.0, 1, .4, 3.0, flow, rho)
    i = 3 + 6 * (3 + 6 * 3)
    assert mc[i] > semi[i] + .01
    assert mc.min() >= 0 and mc.max() <= 1
    for axis in range(3):
        ex, ey = n + (axis == 0), n + (axis == 1)
        src = np.zeros(faces, np.float32)
        src[3 + ex * (3 + ey * 3)] = 1
        out = oracle_face(n, 6.0, axis, 1, .4, 3.0, flow, src)
        assert out.min() >= 0 and out.max() <= 1
        ref = oracle_face(n, 6.0, axis, 0, .4, 3.0, flow, src)
        assert max_abs(oracle_face(n, 6.0, axis, 1, .4, .001, flow, src), ref) <= 1e-5


def test_advection_contracts():
    # AdvectionTests.cpp:159-177 (sizes/aliasing live in the product's ABI)
    n = 3
    cells, faces = sizes(n)
    flow = [np.zeros(faces, np.float32)] * 3
    for dt in (-1.0, float("nan")):
        with pytest.raises(ValueError):
            oracle_scalar(n, 4.5, 0, dt, 3.0, flow, np.zeros(cells, np.float32))
    with pytest.raises(ValueError):
        oracle_face(n, 4.5, 0, 0, .1, 0.0, flow, np.zeros(faces, np.float32))


def test_closed_boundaries_and_flux_sum():
    # NumericsTests.cpp:129-148
    n = 4
    cells, faces = sizes(n)
    u, v, w = np.full(faces, 1, np.float32), np.full(faces, 2, np.float32), np.full(faces, 3, np.float32)
    lib = oracle.oracle_lib()
    lib.orc_enforce_walls(n, ptr(u), ptr(v), ptr(w))
    U = u.reshape(n, n, n + 1)
    assert np.all(U[:, :, 0] == 0) and np.all(U[:, :, n] == 0) and np.all(U[:, :, 1:n] == 1)
    V = v.reshape(n, n + 1, n)
    assert np.all(V[:, 0, :] == 0) and np.all(V[:, n, :] == 0) and np.all(V[:, 1:n, :] == 2)
    W = w.reshape(n + 1, n, n)
    assert np.all(W[0] == 0) and np.all(W[n] == 0) and np.all(W[1:n] == 3)
    div = np.zeros(cells, np.float32)
    lib.orc_divergence(n, 4.5, ptr(u), ptr(v), ptr(w), ptr(div))
    assert abs(float(np.sum(div, dtype=np.float64))) <= 1e-5


def manufactured(n=5):
    """NumericsTests.cpp:150-170: exact x = (1 + x + 2y + 3z)/8 on a floor-6
    system with full diagonal 6 everywhere."""
    x = np.array([(1 + i + 2 * j + 3 * k) * .125 for k in range(n) for j in range(n) for i in range(n)],
                 np.float32)
    return x


def test_sor_solves_neumann_system_to_brute_force_solution():
    # A dense solve of the same 4^3 Neumann system pins the sweep: SOR's
    # converged x must match numpy's least-squares solution (zero mean).
    n = 4
    N = n ** 3
    A = np.zeros((N, N))
    for k in range(n):
        for j in range(n):
            for i in range(n):
                r = i + n * (j + n * k)
                for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < n and 0 <= jj < n and 0 <= kk < n:
                        A[r, r] += 1
                        A[r, ii + n * (jj + n * kk)] = -1
    rng = np.random.default_rng(3)
    b = rng.standard_normal(N).astype(np.float32)
    b -= np.float32(b.astype(np.float64).mean())
    xs = np.linalg.lstsq(A, b.astype(np.float64), rcond=None)[0]
    xs -= xs.mean()
    o = options(max_iterations=3000, fixed_iterations=1, relaxation=1.5, diagonal_floor=0.0)
    x = np.zeros(N, np.float32)
    r = oracle.VgxProjectionResult() if hasattr(oracle, "VgxProjectionResult") else None
    from vapor_amd.abi import VgxProjectionResult

    r = VgxProjectionResult()
    oracle.oracle_lib().orc_sor_solve(n, 1.5, 0.0, ptr(f32(b)), ptr(x), ctypes.byref(o), ctypes.byref(r))
    x = x.astype(np.float64) - x.astype(np.float64).mean()
    assert np.max(np.abs(x - xs)) <= 1e-4 * max(1, np.max(np.abs(xs)))
    # residual function agrees with the dense product
    x32 = f32(x)
    rinf = oracle.oracle_lib().orc_residual_inf(n, ptr(f32(b)), ptr(x32))
    dense = np.max(np.abs(b.astype(np.float64) - A @ x32.astype(np.float64)))
    assert abs(rinf - dense) <= 1e-12 * max(1, dense)


def test_projection_of_gradient_field_recovers_potential():
    # NumericsTests.cpp:255-278: 8^3, box 4, u = grad(phi) projects to zero and
    # p recovers phi - mean(phi).
    n, box, dt = 8, 4.0, 0.1
    cells, faces = sizes(n)
    phi = np.zeros(cells, np.float32)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                phi[x + n * (y + n * z)] = np.float32(np.sin(np.float32(0.3) * x) * np.cos(np.float32(0.4) * y)
                                                      + np.float32(0.2) * z * z / (n * n))
    h = np.float32(box / n)
    factor = np.float32(dt) / h
    u, v, w = (np.zeros(faces, np.float32) for _ in range(3))
    C = lambda x, y, z: x + n * (y + n * z)  # noqa: E731
    for z in range(n):
        for y in range(n):
            for x in range(1, n):
                u[x + (n + 1) * (y + n * z)] = factor * (phi[C(x, y, z)] - phi[C(x - 1, y, z)])
    for z in range(n):
        for y in range(1, n):
            for x in range(n):
                v[x + n * (y + (n + 1) * z)] = factor * (phi[C(x, y, z)] - phi[C(x, y - 1, z)])
    for z in range(1, n):
        for y in range(n):
            for x in range(n):
                w[x + n * (y + n * z)] = factor * (phi[C(x, y, z)] - phi[C(x, y, z - 1)])
    o = options(dt=dt, relaxation=1.5, max_iterations=4000, absolute_tolerance=1e-5, relative_tolerance=1e-6,
                diagonal_floor=0.0)
    (pu, pv, pw), p, r, _ = oracle_project(n, box, o, (u, v, w))
    assert r.converged
    assert max(np.max(np.abs(pu)), np.max(np.abs(pv)), np.max(np.abs(pw))) <= 1e-5
    assert max_abs(p, phi.astype(np.float64) - phi.astype(np.float64).mean()) <= 1e-4


def test_solenoidal_field_is_a_fixed_point():
    # NumericsTests.cpp:311-325
    n = 4
    cells, faces = sizes(n)
    psi = lambda x, y: float(x * (n - x) * y * (n - y))  # noqa: E731
    u, v, w = (np.zeros(faces, np.float32) for _ in range(3))
    for z in range(n):
        for y in range(n):
            for x in range(n + 1):
                u[x + (n + 1) * (y + n * z)] = psi(x, y + 1) - psi(x, y)
    for z in range(n):
        for y in range(n + 1):
            for x in range(n):
                v[x + n * (y + (n + 1) * z)] = -(psi(x + 1, y) - psi(x, y))
    o = options(max_iterations=4000, absolute_tolerance=1e-5, relative_tolerance=1e-6, diagonal_floor=0.0)
    (pu, pv, pw), p, r, _ = oracle_project(n, float(n), o, (u, v, w))
    assert r.converged and r.iterations == 0
    assert np.array_equal(pu, u) and np.array_equal(pv, v) and np.array_equal(pw, w)


@pytest.mark.parametrize("n", [2, 3, 6, 12])
@pytest.mark.parametrize("iters", [1, 8, 64])
def test_floor6_sor_equals_ghost_centre_loop(n, iters):
    # NumericsTests.cpp:341-384: the floor-6 projection equals the legacy
    # "ghost = centre" red-black loop; restated here independently in numpy.
    cells, faces = sizes(n)
    flows = []
    for comp in (1, 2, 3):
        flows.append(np.sin(np.arange(1, faces + 1, dtype=np.float32) * np.float32(0.13 * comp)).astype(np.float32))
    dt, omega = np.float32(1 / 60), np.float32(1.95)
    o = options(dt=float(dt), relaxation=float(omega), max_iterations=iters, fixed_iterations=1, diagonal_floor=6.0)
    (pu, pv, pw), p, r, _ = oracle_project(n, 4.5, o, flows)
    assert r.iterations == iters
    # legacy loop
    u, v, w = (f.copy() for f in flows)
    oracle.oracle_lib().orc_enforce_walls(n, ptr(u), ptr(v), ptr(w))
    h = np.float32(np.float32(4.5) / np.float32(n))
    hInv = np.float32(1) / h
    U, V, W = u.reshape(n, n, n + 1), v.reshape(n, n + 1, n), w.reshape(n + 1, n, n)
    div = ((U[:, :, 1:] - U[:, :, :-1]) + (V[:, 1:, :] - V[:, :-1, :]) + (W[1:] - W[:-1])) * hInv
    P = np.zeros((n, n, n), np.float32)
    k_, j_, i_ = np.meshgrid(range(n), range(n), range(n), indexing="ij")
    for _ in range(iters):
        for parity in (0, 1):
            Pp = np.pad(P, 1, mode="edge")  # ghost = centre at the walls
            s = (Pp[1:-1, 1:-1, :-2] + Pp[1:-1, 1:-1, 2:] + Pp[1:-1, :-2, 1:-1] + Pp[1:-1, 2:, 1:-1]
                 + Pp[:-2, 1:-1, 1:-1] + Pp[2:, 1:-1, 1:-1])
            new = (1 - omega) * P + omega * (s - div * h * h / dt) / 6
            mask = ((i_ + j_ + k_ + parity) & 1) == 0
            P = np.where(mask, new, P).astype(np.float32)
    U[:, :, 1:n] -= dt * (P[:, :, 1:] - P[:, :, :-1]) * hInv
    V[:, 1:n, :] -= dt * (P[:, 1:, :] - P[:, :-1, :]) * hInv
    W[1:n] -= dt * (P[1:] - P[:-1]) * hInv
    oracle.oracle_lib().orc_enforce_walls(n, ptr(u), ptr(v), ptr(w))
    for a, b in ((pu, u), (pv, v), (pw, w)):
        assert max_abs(a, b) <= 2e-5


def test_projection_rest_and_validation():
    # NumericsTests.cpp:236-253
    for n in (1, 5, 2):
        cells, faces = sizes(n)
        flow = [np.zeros(faces, np.float32)] * 3
        (u, v, w), p, r, _ = oracle_project(n, 4.5, options(max_iterations=4000), flow)
        assert r.converged and r.iterations == 0 and r.divergence_after == 0
        for dt in (0.0, float("nan"), np.finfo(np.float32).smallest_subnormal):
            with pytest.raises(ValueError):
                oracle_project(n, 4.5, options(dt=float(dt)), flow)


def test_neumann_matrix_diagonals():
    # NumericsTests.cpp:327-339 through the residual: A 1 = 0 on every row,
    # corner rows have 3 neighbours, interior 6.
    n = 4
    lib = oracle.oracle_lib()
    ones = np.ones(n ** 3, np.float32)
    zero = np.zeros(n ** 3, np.float32)
    assert lib.orc_residual_inf(n, ptr(zero), ptr(ones)) == 0.0
    e = np.zeros(n ** 3, np.float32)
    e[0] = 1
    assert lib.orc_residual_inf(n, ptr(zero), ptr(e)) == 3.0  # corner: |A e|max = diag 3
    e[:] = 0
    e[1 + n * (1 + n * 1)] = 1
    assert lib.orc_residual_inf(n, ptr(zero), ptr(e)) == 6.0
