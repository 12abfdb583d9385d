"""Pins the oracle's signed distance field and Gaussian smoothing
(src/sdf_curvature.rs:162-269) before it checks the device path.

The checks do not restate the oracle's loops: the reference's own unit tests
(sdf_curvature.rs:783-796), scipy's Gaussian filter (the same truncate-4,
"nearest" definition, summed in another order, so within rounding), exact
closed forms (an impulse's response is the product of the weights, rounded in
pass order), and brute-force distances on small volumes.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

import _oracle as orc


# --- gaussian_kernel ------------------------------------------------------------
def test_kernel_is_normalized_known_answer():
    # sdf_curvature.rs:783-789
    for sigma in (0.5, 1.0, 1.5, 2.0):
        k = orc.gaussian_kernel(sigma)
        assert abs(k.sum() - 1.0) < 1e-14


@pytest.mark.parametrize("sigma", [0.3, 0.5, 1.0, 1.5, 2.86, 4.0, 7.3])
def test_kernel_matches_scipy_truncate4(sigma):
    k = orc.gaussian_kernel(sigma)
    r = int(4.0 * sigma + 0.5)
    assert k.size == 2 * r + 1
    impulse = np.zeros(2 * r + 1)
    impulse[r] = 1.0
    ref = ndimage.gaussian_filter1d(impulse, sigma, mode="constant", truncate=4.0)
    np.testing.assert_allclose(k, ref[::-1], rtol=1e-14, atol=1e-17)
    assert np.array_equal(k, k[::-1])  # symmetric to the bit
    assert np.all(np.diff(k[: r + 1]) > 0)


def test_kernel_degenerate_sigmas():
    assert orc.gaussian_kernel(0.0).tolist() == [1.0]
    assert orc.gaussian_kernel(-2.0).tolist() == [1.0]
    assert orc.gaussian_kernel(0.12).tolist() == [1.0]  # r = floor(0.98) = 0
    k = orc.gaussian_kernel(float("nan"))  # `as isize` of NaN is 0: one NaN weight
    assert k.size == 1 and math.isnan(k[0])
    assert orc.gaussian_kernel(0.125).size == 3  # r = floor(1.0) = 1


# --- gaussian_smooth_nearest_in_place ---------------------------------------------
def test_constant_field_known_answer():
    # sdf_curvature.rs:791-796
    out = orc.gaussian_smooth_nearest(np.full(9 * 8 * 7, 3.25), 9, 8, 7, 1.5)
    assert np.all(np.abs(out - 3.25) < 1e-12)


def test_impulse_response_is_the_product_of_weights():
    # an interior impulse: every pass adds exactly one nonzero term, so the
    # output is w[kz] * (w[ky] * w[kx]) rounded in that order, to the bit
    w_, h_, d_ = 23, 19, 17
    sigma = 1.3
    k = orc.gaussian_kernel(sigma)
    r = k.size // 2
    cx, cy, cz = 11, 8, 9
    data = np.zeros(w_ * h_ * d_)
    data[(cz * h_ + cy) * w_ + cx] = 1.0
    out = orc.gaussian_smooth_nearest(data, w_, h_, d_, sigma).reshape(d_, h_, w_)
    want = np.zeros((d_, h_, w_))
    for kz in range(k.size):
        for ky in range(k.size):
            for kx in range(k.size):
                want[cz + kz - r, cy + ky - r, cx + kx - r] = k[kz] * (k[ky] * k[kx])
    assert np.array_equal(out, want)


@pytest.mark.parametrize("shape,sigma", [((7, 9, 11), 1.5), ((4, 30, 5), 2.2), ((13, 6, 8), 0.7)])
def test_matches_scipy_nearest_mode(shape, sigma):
    d_, h_, w_ = shape
    rng = np.random.default_rng(7)
    data = rng.standard_normal(shape)
    out = orc.gaussian_smooth_nearest(data.reshape(-1), w_, h_, d_, sigma).reshape(shape)
    ref = ndimage.gaussian_filter(data, sigma, mode="nearest", truncate=4.0)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-13)


def test_brute_force_order_on_a_tiny_volume():
    # the reference's loop nest written out per output: x, y, z passes, each
    # output summed in tap order with clamped indices (bit-exact)
    w_, h_, d_ = 5, 4, 3
    sigma = 0.9
    k = [float(v) for v in orc.gaussian_kernel(sigma)]
    r = len(k) // 2
    rng = np.random.default_rng(3)
    data = [float(v) for v in rng.standard_normal(w_ * h_ * d_)]

    def idx(x, y, z):
        return (z * h_ + y) * w_ + x

    def cl(v, n):
        return min(max(v, 0), n - 1)

    cur = data
    for axis in range(3):
        nxt = [0.0] * len(cur)
        for z in range(d_):
            for y in range(h_):
                for x in range(w_):
                    acc = 0.0
                    for ki, wk in enumerate(k):
                        o = ki - r
                        if axis == 0:
                            j = idx(cl(x + o, w_), y, z)
                        elif axis == 1:
                            j = idx(x, cl(y + o, h_), z)
                        else:
                            j = idx(x, y, cl(z + o, d_))
                        acc += wk * cur[j]
                    nxt[idx(x, y, z)] = acc
        cur = nxt
    out = orc.gaussian_smooth_nearest(np.array(data), w_, h_, d_, sigma)
    assert out.tolist() == cur


def test_identity_kernels_leave_data():
    data = np.random.default_rng(1).standard_normal(60)
    for s in (0.0, -1.0, 0.1, float("nan")):
        assert np.array_equal(orc.gaussian_smooth_nearest(data, 5, 4, 3, s), data)


# --- signed_distance_field_um --------------------------------------------------------
def _brute_sq(sel):
    sites = np.argwhere(~sel)
    out = np.zeros(sel.shape, np.int64)
    for p in np.argwhere(sel):
        out[tuple(p)] = ((sites - p) ** 2).sum(axis=1).min()
    return out


@pytest.mark.parametrize("seed,shape,frac", [(0, (6, 7, 8), 0.4), (1, (5, 9, 4), 0.7),
                                             (2, (8, 8, 8), 0.1)])
def test_sdf_brute_force(seed, shape, frac):
    rng = np.random.default_rng(seed)
    bone = (rng.random(shape) < frac).astype(np.uint8) * 255
    h = 10.5
    d_, h_, w_ = shape
    phi = orc.signed_distance_field_um(bone.reshape(-1), w_, h_, d_, h).reshape(shape)
    inside = _brute_sq(bone != 0)
    outside = _brute_sq(bone == 0)
    want = np.where(bone != 0, np.sqrt(inside.astype(np.float64)) * h,
                    -np.sqrt(outside.astype(np.float64)) * h)
    assert np.array_equal(phi, want)
    assert np.all(phi[bone != 0] >= h) and np.all(phi[bone == 0] <= -h)


def test_sdf_errors_follow_the_edt_order():
    with pytest.raises(orc.OracleError, match="all-foreground"):
        orc.signed_distance_field_um(np.ones(8, np.uint8), 2, 2, 2, 1.0)
    with pytest.raises(orc.OracleError, match="all-background"):
        orc.signed_distance_field_um(np.zeros(8, np.uint8), 2, 2, 2, 1.0)
