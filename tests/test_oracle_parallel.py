"""The oracle's OpenMP build (used only to time the CPU path on all cores)
steps exactly as the serial checker: its parallel loops are the advection
samples and the colour passes of red-black SOR, whose updates are
independent, and every sum stays sequential."""
import numpy as np

import oracle
from helpers import bits_equal, variant_params


def test_openmp_oracle_is_bit_identical_to_the_serial_oracle():
    p = variant_params(1, 3, sor_omega=1.8, solve_mode=1, tolerance=1e-5, max_sweeps=3000, check_interval=8,
                       warm_start=1, stir_strength=0.5)
    a = oracle.OracleSim(18, 1.0, p)
    b = oracle.OracleSim(18, 1.0, p, parallel=True)
    for _ in range(3):
        a.advance(1 / 60)
        b.advance(1 / 60)
        assert a.last_projection(0).iterations == b.last_projection(0).iterations
    for x, y in zip(a.fields(), b.fields()):
        assert bits_equal(x, y)
    assert np.max(a.fields()[0]) > 0
