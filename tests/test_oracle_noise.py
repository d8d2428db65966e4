"""Counter-based initial velocity noise (extension, no reference code): the
oracle's Philox4x32-10 against the generator's published known-answer
vectors, its log / cos polynomials against numpy in FP64, and the variates'
distribution."""
import ctypes

import numpy as np
import pytest
from scipy import stats

import oracle

U32x4 = ctypes.c_uint32 * 4
U32x2 = ctypes.c_uint32 * 2


def philox(ctr, key):
    out = U32x4()
    oracle.oracle_lib().orc_philox4x32_10(U32x4(*ctr), U32x2(*key), out)
    return list(out)


# Random123's kat_vectors for philox4x32 with 10 rounds (counter, key, output).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox4x32_10_known_answers(ctr, key, want):
    assert philox(ctr, key) == list(want)


def test_noise_log_accuracy():
    lib = oracle.oracle_lib()
    xs = np.concatenate([np.arange(1, 1 << 12, dtype=np.float64) / (1 << 24),
                         np.linspace(1e-4, 1.0, 4001), [0.5, 0.25, 1.0, 0.70710677, 0.7071068]]).astype(np.float32)
    got = np.array([lib.orc_noise_log(float(x)) for x in xs], np.float64)
    want = np.log(xs.astype(np.float64))
    err = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    assert err.max() < 3e-7
    assert lib.orc_noise_log(1.0) == 0.0


def test_noise_cos_accuracy():
    lib = oracle.oracle_lib()
    ts = np.concatenate([np.linspace(0, 1, 8001, endpoint=False), [0.25, 0.5, 0.75, 0.125]]).astype(np.float32)
    got = np.array([lib.orc_noise_cos2pi(float(t)) for t in ts], np.float64)
    want = np.cos(2 * np.pi * ts.astype(np.float64))
    assert np.abs(got - want).max() < 3e-7
    assert lib.orc_noise_cos2pi(0.0) == 1.0


def test_noise_is_standard_normal():
    n = 40
    cells, faces = oracle.sizes(n)
    u, v, w = (np.zeros(faces, np.float32) for _ in range(3))
    oracle.oracle_lib().orc_add_velocity_noise(n, 1.0, 12345, oracle.ptr(u), oracle.ptr(v), oracle.ptr(w))
    z = np.concatenate([u, v, w]).astype(np.float64)
    assert abs(z.mean()) < 5 / np.sqrt(z.size)
    assert abs(z.std() - 1.0) < 5e-3
    assert stats.kstest(z, "norm").pvalue > 1e-3
    # streams and entries are independent: no correlation between neighbours or axes
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.01
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.01


def test_noise_sigma_scales_and_seed_changes():
    n = 8
    cells, faces = oracle.sizes(n)
    a = [np.zeros(faces, np.float32) for _ in range(3)]
    b = [np.zeros(faces, np.float32) for _ in range(3)]
    c = [np.zeros(faces, np.float32) for _ in range(3)]
    lib = oracle.oracle_lib()
    lib.orc_add_velocity_noise(n, 1.0, 7, *map(oracle.ptr, a))
    lib.orc_add_velocity_noise(n, 0.05, 7, *map(oracle.ptr, b))
    lib.orc_add_velocity_noise(n, 1.0, 8, *map(oracle.ptr, c))
    for x, y in zip(a, b):
        assert np.array_equal(y, (np.float32(0.05) * x).astype(np.float32))
    assert not np.array_equal(a[0], c[0])
    # the variate of entry i is the one a single call returns
    assert a[1][5] == lib.orc_noise_normal(7, 1, 5)
