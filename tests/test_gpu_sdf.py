"""GPU parity of the signed distance field and its Gaussian smoothing
(src/sdf_curvature.rs:162-269) through the C ABI.  Every value is f64 and must
be bit-identical to the oracle's: the square root is correctly rounded and the
smoothing adds its products one by one in tap order, as the reference does."""
import numpy as np
import pytest

import _oracle as orc
import bone_morphometry_amd as bm
from bone_morphometry_amd import synth, workloads

pytestmark = pytest.mark.gpu


def vol(a, w, h, d):
    return bm.BinaryVolume(np.asarray(a, np.uint8).reshape(-1), w, h, d)


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_diff(a, b):
    i = np.flatnonzero(a.view(np.uint64) != b.view(np.uint64))
    return f"{i.size} differ, first at {i[:3]}: {a[i[:3]]} vs {b[i[:3]]}" if i.size else ""


CASES = [  # (w, h, d, gaussian sigma, bv/tv, seed, label)
    (40, 33, 29, 2.0, 0.3, 21, "ragged"),
    (64, 64, 64, 3.0, 0.25, 22, "cube64"),
    (130, 70, 45, 2.5, 0.15, 23, "wide"),
    (17, 150, 90, 4.0, 0.4, 24, "tall"),
]


@pytest.mark.parametrize("w,h,d,sigma,frac,seed,label", CASES)
def test_sdf_matches_oracle(w, h, d, sigma, frac, seed, label):
    bone = synth.volume(w, h, d, sigma, frac, seed)
    hs = 10.5
    got = bm.signed_distance_field_um(vol(bone, w, h, d), hs)
    want = orc.signed_distance_field_um(bone, w, h, d, hs)
    assert same_bits(got, want), first_diff(got, want)


def test_sdf_square_roots_over_a_wide_range():
    # one background voxel in a corner: d2 runs over every sum of three squares
    # up to 3 * 159^2, so the device square root is checked on ~10^5 values
    n = 160
    bone = np.full(n ** 3, 255, np.uint8)
    bone[0] = 0
    hs = 7.25
    got = bm.signed_distance_field_um(vol(bone, n, n, n), hs)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d2 = (x * x + y * y + z * z).reshape(-1).astype(np.float64)
    want = np.sqrt(d2) * hs
    want[0] = -np.sqrt(1.0) * hs
    assert same_bits(got, want), first_diff(got, want)


def test_sdf_errors():
    with pytest.raises(bm.BmError, match="all-foreground"):
        bm.signed_distance_field_um(vol(np.ones(27), 3, 3, 3), 1.0)
    with pytest.raises(bm.BmError, match="all-background"):
        bm.signed_distance_field_um(vol(np.zeros(27), 3, 3, 3), 1.0)
    with pytest.raises(bm.BmError, match="all-background"):
        bm.smoothed_signed_distance_field_um(vol(np.zeros(27), 3, 3, 3), 1.0, 2.0)


SMOOTH_CASES = [  # (w, h, d, sigma_vox): register-block path, short kernels, big LDS, no LDS
    (9, 8, 7, 1.5),          # the reference's own test shape
    (40, 33, 29, 0.6),       # r = 2: taps fewer than a register block
    (65, 64, 63, 2.857142857142857),  # 30 um at 10.5 um voxels: r = 11
    (130, 20, 70, 3.0),
    (5, 3, 200, 4.4),        # axes shorter than the radius
    (1, 1, 50, 2.0),         # degenerate axes
    (48, 40, 36, 20.0),      # r = 80: a tile of 116 KiB of LDS
    (40, 30, 20, 31.0),      # r = 124: beyond LDS, taps read from memory
]


@pytest.mark.parametrize("w,h,d,sigma", SMOOTH_CASES)
def test_smoothing_matches_oracle(w, h, d, sigma):
    rng = np.random.default_rng(w * 131 + h * 17 + d)
    data = rng.standard_normal(w * h * d) * 50.0
    want = orc.gaussian_smooth_nearest(data, w, h, d, sigma)
    got = data.copy()
    bm.gaussian_smooth_nearest_in_place(got, w, h, d, sigma)
    assert same_bits(got, want), first_diff(got, want)


def test_smoothing_constant_field_known_answer():
    # sdf_curvature.rs:791-796
    data = np.full(9 * 8 * 7, 3.25)
    bm.gaussian_smooth_nearest_in_place(data, 9, 8, 7, 1.5)
    assert np.all(np.abs(data - 3.25) < 1e-12)


def test_smoothing_non_finite_values_stay_local():
    w, h, d = 30, 20, 25
    rng = np.random.default_rng(5)
    data = rng.standard_normal(w * h * d)
    data[[17, 4000, 9000]] = [np.nan, np.inf, -np.inf]
    want = orc.gaussian_smooth_nearest(data, w, h, d, 1.7)
    got = data.copy()
    bm.gaussian_smooth_nearest_in_place(got, w, h, d, 1.7)
    assert same_bits(got, want), first_diff(got, want)


def test_smoothing_identity_kernels_and_errors():
    data = np.random.default_rng(2).standard_normal(60)
    for s in (0.0, -3.0, 0.1, float("nan")):
        got = data.copy()
        bm.gaussian_smooth_nearest_in_place(got, 5, 4, 3, s)
        assert same_bits(got, data)
    with pytest.raises(bm.BmError, match="data holds 60 values, the volume has 61 voxels"):
        bm.gaussian_smooth_nearest_in_place(data.copy(), 61, 1, 1, 1.0)


@pytest.mark.parametrize("w,h,d,sigma,frac,seed,label", CASES)
def test_smoothed_sdf_matches_oracle_and_separate_calls(w, h, d, sigma, frac, seed, label):
    bone = synth.volume(w, h, d, sigma, frac, seed)
    hs, sigma_um = 10.5, 30.0
    sv = sigma_um / hs
    got = bm.smoothed_signed_distance_field_um(vol(bone, w, h, d), hs, sv)
    want = orc.gaussian_smooth_nearest(orc.signed_distance_field_um(bone, w, h, d, hs), w, h, d, sv)
    assert same_bits(got, want), first_diff(got, want)
    two = bm.signed_distance_field_um(vol(bone, w, h, d), hs)
    bm.gaussian_smooth_nearest_in_place(two, w, h, d, sv)
    assert same_bits(got, two)


def test_smoothed_sdf_identity_kernel_is_the_field():
    w, h, d = 30, 31, 32
    bone = synth.volume(w, h, d, 2.0, 0.3, 8)
    a = bm.smoothed_signed_distance_field_um(vol(bone, w, h, d), 2.0, 0.05)
    assert same_bits(a, bm.signed_distance_field_um(vol(bone, w, h, d), 2.0))


def test_device_resident_forms():
    from bone_morphometry_amd import hip
    w, h, d = 70, 50, 40
    n = w * h * d
    bone = synth.volume(w, h, d, 2.5, 0.3, 31)
    hs, sv = 10.0, 3.0
    want_sdf = orc.signed_distance_field_um(bone, w, h, d, hs)
    want = orc.gaussian_smooth_nearest(want_sdf, w, h, d, sv)
    d_bone, d_out = hip.DeviceBuffer(n), hip.DeviceBuffer(n * 8)
    d_bone.upload(bone)
    got = np.empty(n, np.float64)
    bm.signed_distance_field_um_device(d_bone.value, w, h, d, hs, d_out.value)
    d_out.download(got)
    assert same_bits(got, want_sdf)
    bm.gaussian_smooth_nearest_in_place_device(d_out.value, w, h, d, sv)
    d_out.download(got)
    assert same_bits(got, want)
    d_out.upload(np.zeros(n))
    stream = hip.stream()
    bm.smoothed_signed_distance_field_um_device(d_bone.value, w, h, d, hs, sv, d_out.value,
                                                stream=stream.value)
    d_out.download(got)
    assert same_bits(got, want)


def _exact_smoothed_at(phi, shape, k, points):
    """The reference's three passes evaluated at a few voxels only, in numpy:
    each partial sum adds one rounded product at a time, in tap order."""
    d_, h_, w_ = shape
    r = k.size // 2
    f = phi.reshape(shape)
    offs = np.arange(k.size) - r
    out = []
    for (x, y, z) in points:
        zs = np.clip(z + offs, 0, d_ - 1)
        ys = np.clip(y + offs, 0, h_ - 1)
        xs = np.clip(x + offs, 0, w_ - 1)
        block = f[np.ix_(zs, ys, xs)]           # [kz, ky, kx]
        acc_x = np.zeros(block.shape[:2])
        for kx in range(k.size):
            acc_x = acc_x + k[kx] * block[:, :, kx]
        acc_y = np.zeros(block.shape[0])
        for ky in range(k.size):
            acc_y = acc_y + k[ky] * acc_x[:, ky]
        acc = 0.0
        for kz in range(k.size):
            acc = acc + k[kz] * acc_y[kz]
        out.append(acc)
    return np.array(out)


def test_full_size_512_properties():
    # the 512^3 trabecular workload: fused path == field then smoothing, and
    # 400 voxels (edges and corners included) against an exact evaluation
    wl = workloads.plain_512()
    w, h, d = wl.width, wl.height, wl.depth
    hs = wl.spacing[0]
    sv = 30.0 / hs
    v = vol(wl.bone, w, h, d)
    phi = bm.signed_distance_field_um(v, hs)
    inside = wl.bone != 0
    assert np.all(phi[inside] >= hs) and np.all(phi[~inside] <= -hs)
    fused = bm.smoothed_signed_distance_field_um(v, hs, sv)
    two = phi.copy()
    bm.gaussian_smooth_nearest_in_place(two, w, h, d, sv)
    assert same_bits(fused, two), first_diff(fused, two)
    k = bm.gaussian_kernel(sv)
    rng = np.random.default_rng(9)
    pts = [tuple(int(c) for c in rng.integers(0, 512, 3)) for _ in range(392)]
    pts += [(0, 0, 0), (511, 511, 511), (0, 511, 3), (5, 0, 511), (511, 2, 0), (3, 4, 5),
            (508, 507, 506), (255, 256, 257)]
    want = _exact_smoothed_at(phi, (d, h, w), k, pts)
    got = np.array([fused[(z * h + y) * w + x] for (x, y, z) in pts])
    assert same_bits(got, want), first_diff(got, want)
