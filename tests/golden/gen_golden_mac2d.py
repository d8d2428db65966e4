"""Writes the MAC 2D golden fixtures in tests/golden/ from vapor's own 2D solvers.

    make -C oracle ref && python tests/golden/gen_golden_mac2d.py

The reference is the one built by oracle/Makefile from the vapor sources
(oracle/_ref/libvapor_ref2d.so); this script only calls it through its C shim
and stores seeded inputs and the outputs as data.

Fixtures:
  mac2d_stages.npz        SemiLagrangian2D::advect per field and dt (n = 9),
                          CpuPressureSolver2D::project on random and
                          manufactured faces (n = 1, 8, 15; several boxes,
                          dt, fixed and tolerance budgets): faces, pressure,
                          result records
  mac2d_steps.npz         MacGridFluidSolver2D trajectories (simple and
                          midpoint reflection; n = 16 box 1 and n = 24 box 4)
  mac2d_hist_steps.npz    ReflectionMacFluidSolver2D trajectories (semi-
                          Lagrangian and MacCormack smoke, stirring off and on)
"""
from __future__ import annotations

import ctypes
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import mac2d as O  # noqa: E402
from vapor_amd.abi import VgxProjectionOptions, VgxProjectionResult, Vgx2Params, Vgx2rParams  # noqa: E402

RESULT_FIELDS = ("iterations", "converged", "initial_residual", "final_residual", "residual_available",
                 "divergence_before", "divergence_after", "removed_rhs_mean")


def result_row(r: VgxProjectionResult) -> np.ndarray:
    return np.array([getattr(r, f) for f in RESULT_FIELDS], np.float64)


def params2d(**kw) -> Vgx2Params:
    p = Vgx2Params()
    O.ref2d_lib().r2_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def options(dt, relaxation, max_iterations, fixed=0, check=8, abs_tol=1e-6, rel_tol=1e-5) -> VgxProjectionOptions:
    o = VgxProjectionOptions()
    o.dt, o.relaxation, o.max_iterations, o.residual_check_interval = dt, relaxation, max_iterations, check
    o.absolute_tolerance, o.relative_tolerance, o.fixed_iterations = abs_tol, rel_tol, fixed
    return o


def stages():
    lib = O.ref2d_lib()
    rng = np.random.default_rng(2024)
    out = {}
    n, box = 9, 1.0
    N, F = O.sizes(n)
    h = box / n
    u = (rng.standard_normal(F) * h * 2.0).astype(np.float32)
    v = (rng.standard_normal(F) * h * 2.0).astype(np.float32)
    out["adv_u"], out["adv_v"] = u, v
    for field, count in ((0, N), (1, F), (2, F)):
        src = rng.random(count).astype(np.float32)
        out[f"adv_src{field}"] = src
        for k, dt in enumerate((0.0, 0.125, 0.5, 3.0)):
            o = np.empty(count, np.float32)
            rc = lib.r2_advect(n, box, field, dt, O.fptr(u), O.fptr(v), O.fptr(src), O.fptr(o))
            assert rc == 0
            out[f"adv_out{field}_{k}"] = o
    out["adv_dts"] = np.array([0.0, 0.125, 0.5, 3.0], np.float32)
    cases = []
    for n, box, dt, relax, iters, fixed, kind in ((1, 1.0, .01, 1.6, 4000, 0, "random"),
                                                  (8, 1.0, .01, 1.6, 4000, 0, "gradient"),
                                                  (8, 3.0, .1, 1.6, 4000, 0, "gradient"),
                                                  (15, 1.0, .1, 1.6, 4000, 0, "gradient"),
                                                  (15, 3.0, .01, 1.9, 600, 0, "random"),
                                                  (12, 2.0, .02, 1.6, 17, 1, "random"),
                                                  (16, 4.0, 1 / 60, 1.9, 600, 0, "random"),
                                                  (8, 4.0, 1 / 60, 1.6, 0, 0, "random")):
        N, F = O.sizes(n)
        hh = np.float32(box) / np.float32(n)
        if kind == "gradient":
            # CpuPressureSolver2D's manufactured case (Mac2DTests.cpp:93-107 shape)
            x = (np.arange(n) + .5) / n
            exact = (np.cos(np.pi * x)[None, :] * np.cos(2 * np.pi * x)[:, None]).astype(np.float32).ravel()
            exact = (exact - np.float32(exact.astype(np.float64).mean())).astype(np.float32)
            uu = np.zeros(F, np.float32)
            vv = np.zeros(F, np.float32)
            e = exact.reshape(n, n)
            uu.reshape(n, n + 1)[:, 1:n] = np.float32(dt) / hh * (e[:, 1:] - e[:, :-1])
            vv.reshape(n + 1, n)[1:n, :] = np.float32(dt) / hh * (e[1:, :] - e[:-1, :])
        else:
            uu = (rng.standard_normal(F)).astype(np.float32)
            vv = (rng.standard_normal(F)).astype(np.float32)
        u0, v0 = uu.copy(), vv.copy()
        p = np.zeros(N, np.float64)
        r = VgxProjectionResult()
        o = options(dt, relax, iters, fixed)
        rc = lib.r2_project(n, box, ctypes.byref(o), O.fptr(uu), O.fptr(vv), O.dptr(p), ctypes.byref(r))
        assert rc == 0
        cases.append((n, box, dt, relax, iters, fixed))
        k = len(cases) - 1
        out[f"proj{k}_u0"], out[f"proj{k}_v0"] = u0, v0
        out[f"proj{k}_u"], out[f"proj{k}_v"], out[f"proj{k}_p"] = uu, vv, p
        out[f"proj{k}_result"] = result_row(r)
    out["proj_cases"] = np.array(cases, np.float64)
    np.savez_compressed(HERE / "mac2d_stages.npz", **out)


def steps():
    out = {}
    runs = [(16, 1.0, 0, 8), (16, 1.0, 1, 8), (24, 4.0, 0, 6), (24, 4.0, 1, 6)]
    for k, (n, box, refl, count) in enumerate(runs):
        prm = params2d(reflection=refl)
        sim = O.RefSim2D(n, box, prm)
        for s in range(count):
            assert sim.advance(1 / 60) == 0
            rho, u, v = sim.fields()
            out[f"run{k}_step{s}_rho"], out[f"run{k}_step{s}_u"], out[f"run{k}_step{s}_v"] = rho, u, v
            out[f"run{k}_step{s}_last"] = result_row(sim.projection())
            out[f"run{k}_step{s}_mid"] = result_row(sim.projection(True))
        rc, d = sim.diagnostics()
        out[f"run{k}_diag"] = np.array([rc, d.kinetic_energy, d.max_velocity, d.max_divergence], np.float64)
    out["runs"] = np.array(runs, np.float64)
    np.savez_compressed(HERE / "mac2d_steps.npz", **out)


def hist_steps():
    out = {}
    runs = [(16, 0, 0.0, 10), (16, 1, 0.0, 10), (24, 1, 0.7, 8), (24, 0, 0.5, 8)]
    for k, (n, mc, stir, count) in enumerate(runs):
        prm = Vgx2rParams()
        O.ref2d_lib().r2h_params(n, -1, ctypes.byref(prm))
        prm.maccormack_smoke, prm.stir_strength = mc, stir
        sim = O.RefHist2D(n, prm)
        for s in range(count):
            assert sim.advance(1 / 60) == 0
            f = sim.fields()
            for name in ("smoke", "vx", "vy", "pressure"):
                out[f"run{k}_step{s}_{name}"] = f[name]
            out[f"run{k}_step{s}_random"] = np.array([sim.random_state()], np.uint32)
        d = sim.diagnostics()
        out[f"run{k}_diag"] = np.array([d.kinetic_energy, d.max_velocity, d.max_divergence], np.float64)
    out["runs"] = np.array(runs, np.float64)
    np.savez_compressed(HERE / "mac2d_hist_steps.npz", **out)


if __name__ == "__main__":
    stages()
    steps()
    hist_steps()
    for name in ("mac2d_stages.npz", "mac2d_steps.npz", "mac2d_hist_steps.npz"):
        print(name, (HERE / name).stat().st_size, "bytes")
