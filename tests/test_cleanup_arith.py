"""Floating-point identities the surface clean-up kernel relies on
(k_cleanup_stream in bm_tiles.hip), checked on the CPU with numpy.

A surface voxel's value is the f32 mean of its contributing neighbours:
sum / count, the f32 sum of up to 26 contributions (each a diameter, at least
2) divided by their count (src/local_thickness.rs:243-277).  The kernel
computes it as the f64 product of the sum with 1/count rounded to f64, then
rounds to f32; the quotient is never an f32 rounding midpoint and lies at
least 2^-25/26 (relative) from one, so both roundings agree.  A contribution
is +0.0f or at least 2, so bit 30 of its pattern says whether it counts."""
import numpy as np
import pytest

RCP = np.array([0.0] + [1.0 / n for n in range(1, 27)], np.float64)


def _kernel_mean(s, n):
    return (s.astype(np.float64) * RCP[n]).astype(np.float32)


def _ieee_mean(s, n):
    return s / n.astype(np.float32)  # f32 division, correctly rounded


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_f64_reciprocal_mean_equals_f32_division_random(seed):
    rng = np.random.default_rng(seed)
    # sums of 1..26 diameters (each >= 2): from 2 to a large multiple, across binades
    s = (2.0 ** rng.uniform(1, 40, 2_000_000)).astype(np.float32)
    n = rng.integers(1, 27, s.size)
    np.testing.assert_array_equal(_kernel_mean(s, n), _ieee_mean(s, n))


def test_f64_reciprocal_mean_equals_f32_division_dense():
    # every f32 in [2, 4) and [1024, 1032) with stride 7 ulps, against every count
    a = np.arange(0x40000000, 0x40800000, 7, dtype=np.uint32).view(np.float32)
    b = np.arange(0x44800000, 0x44810000, 1, dtype=np.uint32).view(np.float32)
    for s in (a, b):
        for n in range(1, 27):
            nn = np.full(s.size, n)
            np.testing.assert_array_equal(_kernel_mean(s, nn), _ieee_mean(s, nn), err_msg=f"n={n}")


def test_f64_reciprocal_mean_on_sums_of_diameters():
    # sums built the kernel's way: sequential f32 additions of diameters 2 sqrt(m)
    rng = np.random.default_rng(7)
    m = rng.integers(1, 20000, (200_000, 26))
    d = (2.0 * np.sqrt(m.astype(np.float32))).astype(np.float32)
    n = rng.integers(1, 27, m.shape[0])
    s = np.zeros(m.shape[0], np.float32)
    for k in range(26):
        s = np.where(k < n, s + d[:, k], s).astype(np.float32)
    np.testing.assert_array_equal(_kernel_mean(s, n), _ieee_mean(s, n))


def test_contribution_bit_30():
    finite = np.array([2.0, 2.0000002, 3.999, 4.0, 1e6, 3.4e38], np.float32)
    assert ((finite.view(np.uint32) >> 30) == 1).all()
    assert (np.array([0.0], np.float32).view(np.uint32) >> 30 == 0).all()
    # diameters of every r^2 up to 2^23
    m = np.arange(1, 1 << 23, dtype=np.float64)
    d = (2.0 * np.sqrt(m)).astype(np.float32)
    assert ((d.view(np.uint32) >> 30) == 1).all()


def test_planted_reciprocal_error_is_caught():
    """A reciprocal rounded to f32 instead of f64 gives different means: the
    random test above would see it."""
    rng = np.random.default_rng(3)
    s = (2.0 ** rng.uniform(1, 40, 200_000)).astype(np.float32)
    n = rng.integers(1, 27, s.size)
    bad = (s.astype(np.float64) * RCP.astype(np.float32).astype(np.float64)[n]).astype(np.float32)
    assert (bad != _ieee_mean(s, n)).any()
