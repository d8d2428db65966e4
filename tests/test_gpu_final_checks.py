"""Final residual checks of a tolerance solve (vgx_params.final_check_interval).

With final_check_interval = k, a single-scene tolerance solve checks its
residual every k sweeps once the residual's factor per sweep says the next
check interval reaches the tolerance, instead of once per interval: the solve
stops fewer sweeps past the tolerance. The default (0) keeps one check per
interval, the schedule whose sweep counts equal the oracle's (tested in
test_gpu_parity.py and test_gpu_configs.py).

Two simulations of the same scene step in lockstep, one per schedule: every
projection of both meets the tolerance; a projection with final checks never
takes more than one check interval beyond the fixed schedule's; over the run
the final checks take fewer sweeps; and after the first step the fields agree to
the precision two converged solves allow."""
import numpy as np
import pytest

from vapor_amd import mac3d
from vapor_amd.train_warm_start import scene

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.mark.parametrize("n,seed,steps", [(64, 31, 12), (96, 32, 6)])
def test_final_checks_stop_closer_to_the_tolerance(n, seed, steps):
    base = scene(seed, mac3d)
    tol, interval = base.tolerance, base.check_interval
    fixed = mac3d.GpuMacSimulation3D(n, 1.0, base.copy().update(final_check_interval=0))
    final = mac3d.GpuMacSimulation3D(n, 1.0, base.copy().update(final_check_interval=8))
    totals = {"fixed": 0, "final": 0}
    for step in range(steps):
        fixed.advance(DT)
        final.advance(DT)
        if step == 0:
            # after one step the two solves differ by their stopping points only
            x, y = fixed.state().density(), final.state().density()
            scale = max(float(np.max(np.abs(x))), 1e-30)
            assert float(np.max(np.abs(x - y))) <= 1e-4 * scale
        for k in (1, 0):
            a, b = fixed.lastProjection(k), final.lastProjection(k)
            assert a.converged and b.converged, (step, k)
            assert a.final_residual_l2 <= tol and b.final_residual_l2 <= tol, (step, k)
            assert b.iterations <= a.iterations + interval, (step, k, a.iterations, b.iterations)
            totals["fixed"] += a.iterations
            totals["final"] += b.iterations
    print(f"[gpu final checks] n={n} seed={seed}: {totals['fixed']} sweeps with a check every {interval}, "
          f"{totals['final']} with final checks every 8", flush=True)
    assert totals["final"] < totals["fixed"], totals
    fixed.close()
    final.close()
