"""GPU parity: the HIP path (through the C ABI) against the oracle, the
golden fixtures and the reference's own known answers.  Integer products (the
squared EDT, ridge flags, propagated r^2) and the f32 thickness maps must be
bit-identical; the f64 summaries within 1e-12 relative (reduction order)."""
import math
import os

import numpy as np
import pytest

import _oracle as orc
import bone_morphometry_amd as bm
from _stats import check_stats
from bone_morphometry_amd import synth, workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "edt_scipy.npz")
REL = 1e-12


def vol(a, w, h, d):
    return bm.BinaryVolume(np.asarray(a, np.uint8).reshape(-1), w, h, d)


def _structure(seed, shape, frac):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape)
    for ax in range(3):
        f = (f + np.roll(f, 1, ax) + np.roll(f, -1, ax)) / 3
    return (f > np.quantile(f, 1.0 - frac)).astype(np.uint8)


# --- the reference's unit tests through the C ABI -----------------------------
def test_phase_inside_known_answer():
    bone = vol([0, 1, 0, 1, 0, 1, 0, 1], 2, 2, 2)
    roi = vol([0, 0, 1, 1, 1, 1, 0, 0], 2, 2, 2)
    assert bone.phase_inside(roi, True).data.tolist() == [0, 0, 0, 1, 0, 1, 0, 0]
    assert bone.phase_inside(roi, False).data.tolist() == [0, 0, 1, 0, 1, 0, 0, 0]


def test_edt_known_answers():
    a = np.ones(27, np.uint8)
    a[13] = 0
    e = bm.euclidean_distance_transform_squared(vol(a, 3, 3, 3), bm.Phase.Foreground)
    assert (e[13], e[12], e[0]) == (0, 1, 3)
    b = np.zeros(27, np.uint8)
    b[13] = 1
    e = bm.euclidean_distance_transform_squared(vol(b, 3, 3, 3), bm.Phase.Background)
    assert (e[13], e[12], e[0]) == (0, 1, 3)


def test_edt_all_selected_error():
    with pytest.raises(bm.BmError, match="distance transform is undefined for an "
                                          "all-foreground volume unless an exterior "
                                          "convention is specified"):
        bm.euclidean_distance_transform_squared(vol(np.ones(8), 2, 2, 2), bm.Phase.Foreground)
    with pytest.raises(bm.BmError, match="all-background"):
        bm.local_thickness(vol(np.zeros(8), 2, 2, 2), bm.Phase.Background)


def test_isolated_voxel_has_diameter_two():
    a = np.zeros(27, np.uint8)
    a[13] = 1
    t = bm.local_thickness(vol(a, 3, 3, 3), bm.Phase.Foreground)
    assert abs(t[13] - 2.0) < 1e-6 and (t > 0).sum() == 1


def test_background_matches_complemented_foreground():
    a = np.zeros((5, 5, 5), np.uint8)
    a[1:4, 1:4, 1:4] = 1
    d = bm.local_thickness(vol(a, 5, 5, 5), bm.Phase.Background)
    c = bm.local_thickness(vol(a == 0, 5, 5, 5), bm.Phase.Foreground)
    np.testing.assert_array_equal(d, c)


# --- EDT ------------------------------------------------------------------------
def test_edt_matches_golden_fixtures():
    g = np.load(GOLDEN)
    k = 0
    while f"vol_{k}" in g:
        w, h, d = (int(v) for v in g[f"shape_{k}"])
        got = bm.euclidean_distance_transform_squared(vol(g[f"vol_{k}"], w, h, d),
                                                      bm.Phase.Foreground)
        np.testing.assert_array_equal(got, g[f"d2_{k}"], err_msg=f"case {k}: {w}x{h}x{d}")
        k += 1


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 5, 7), (64, 64, 64), (65, 63, 67),
                                   (600, 5, 4), (5, 700, 3), (4, 3, 1300), (97, 33, 530)])
@pytest.mark.parametrize("phase", [bm.Phase.Foreground, bm.Phase.Background])
def test_edt_matches_oracle(shape, phase):
    w, h, d = shape
    a = _structure(sum(shape) + int(phase), (d, h, w), 0.6)
    if a.all() or not a.any():
        a.flat[0] ^= 1
    got = bm.euclidean_distance_transform_squared(vol(a, w, h, d), phase)
    np.testing.assert_array_equal(got, orc.edt_squared(a, w, h, d, int(phase)))


def test_edt_sparse_sites_long_searches():
    w, h, d = 70, 80, 90
    a = np.ones((d, h, w), np.uint8)
    a[3, 77, 5] = 0
    a[60, 1, 66] = 0
    got = bm.euclidean_distance_transform_squared(vol(a, w, h, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got, orc.edt_squared(a, w, h, d, 0))


@pytest.mark.parametrize("shape,sites", [
    ((600, 5, 4), [(0, 0, 0)]),
    ((5, 700, 3), [(4, 0, 2)]),
    ((4, 3, 1300), [(0, 2, 0)]),
    ((70, 620, 9), [(0, 0, 0), (69, 300, 8)]),
    ((9, 40, 900), [(4, 20, 450), (0, 0, 0)]),
])
def test_edt_distances_beyond_16_bits(shape, sites):
    """Squared distances of 2^16 and more next to smaller ones in the same
    wavefront: the Y and Z passes search in saturated 16-bit pairs and search
    a group again in 32 bits when one of its rows saturates."""
    w, h, d = shape
    a = np.ones((d, h, w), np.uint8)
    for x, y, z in sites:
        a[z, y, x] = 0
    ref = orc.edt_squared(a, w, h, d, 0)
    assert ref.max() >= 1 << 16 and (ref < 1 << 16).sum() > 0
    got = bm.euclidean_distance_transform_squared(vol(a, w, h, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("tiles", ["40,7,41,7", "41,8,36,9", "16,0,18,1", "24,2,512,64"])
@pytest.mark.parametrize("phase", [bm.Phase.Foreground, bm.Phase.Background])
def test_edt_tile_shapes_match_oracle(tiles, phase, monkeypatch):
    """Y and Z tiles of odd halos and odd row counts (chunks rounded down to
    even rows, tiles starting one row early) against the oracle."""
    monkeypatch.setenv("BM_EDT_TILES", tiles)
    w, h, d = 37, 150, 140
    a = _structure(11 + int(phase), (d, h, w), 0.5)
    got = bm.euclidean_distance_transform_squared(vol(a, w, h, d), phase)
    np.testing.assert_array_equal(got, orc.edt_squared(a, w, h, d, int(phase)))


def test_edt_tile_without_chunk_raises(monkeypatch):
    monkeypatch.setenv("BM_EDT_TILES", "21,10,128,16")
    a = _structure(3, (40, 60, 20), 0.5)
    with pytest.raises(bm.BmError, match="y tile chunk below 2 rows"):
        bm.euclidean_distance_transform_squared(vol(a, 20, 60, 40), bm.Phase.Foreground)


# --- local thickness --------------------------------------------------------------
@pytest.mark.parametrize("seed,shape,frac", [(0, (20, 21, 22), 0.3), (1, (33, 17, 25), 0.7),
                                             (2, (40, 40, 40), 0.85), (3, (7, 60, 9), 0.5)])
def test_local_thickness_stages_match_oracle(seed, shape, frac):
    w, h, d = shape
    a = _structure(seed, (d, h, w), frac)
    got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
    ref = orc.local_thickness_stages(a, w, h, d, 0)
    np.testing.assert_array_equal(got["squared_edt"], ref["squared_edt"])
    np.testing.assert_array_equal(got["ridge_flags"], ref["ridge_flags"])
    assert got["ridge_count"] == ref["ridge_count"]
    assert got["fill_visits"] == ref["fill_visits"]
    np.testing.assert_array_equal(got["propagated"], ref["propagated"])
    np.testing.assert_array_equal(got["thickness"], ref["thickness"])


@pytest.mark.parametrize("cap", ["64", "257"])
def test_fill_rounds_match_oracle(cap, monkeypatch):
    """A small LDS budget forces the fill through many radius rounds per tile
    (and buckets larger than the budget through the exact overflow kernel)."""
    monkeypatch.setenv("BM_FILL_CAP", cap)
    w, h, d = 48, 40, 44
    a = _structure(9, (d, h, w), 0.8)
    got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
    ref = orc.local_thickness_stages(a, w, h, d, 0)
    np.testing.assert_array_equal(got["propagated"], ref["propagated"])
    np.testing.assert_array_equal(got["thickness"], ref["thickness"])


@pytest.mark.parametrize("sites,seed", [(24, 21), (60, 22)])
def test_fill_mixed_radius_buckets_match_oracle(sites, seed):
    """Sparse sites give squared distances above 511, so the fill's radius
    buckets hold several r^2 values each: a voxel covered by a ball of a
    bucket must still take a larger ball of the same bucket.  Run twice: the
    order inside a bucket varies between runs, the result may not."""
    rng = np.random.default_rng(seed)
    w, h, d = 72, 64, 56
    a = np.ones(w * h * d, np.uint8)
    a[rng.choice(a.size, sites, replace=False)] = 0
    ref = orc.local_thickness_stages(a, w, h, d, 0)
    assert ref["squared_edt"].max() >= 512
    for _ in range(2):
        got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
        np.testing.assert_array_equal(got["propagated"], ref["propagated"])
        np.testing.assert_array_equal(got["thickness"], ref["thickness"])


@pytest.mark.parametrize("tiles", ["1", "2"])
@pytest.mark.parametrize("radius", [64, 65])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fill_radius_limit_of_dot_form(axis, radius, tiles, monkeypatch):
    """Two planes of sites 2R apart along one axis: the largest radius is
    exactly R.  R = 64 is the largest radius of the fill's dot-product test
    (candidate offsets from the block origin span -64 .. 8 * tiles - 1 + 64);
    R = 65 takes the packed-position test.  One- and two-tile fill blocks."""
    monkeypatch.setenv("BM_FILL_TILES", tiles)
    dims = [11, 10, 9]  # (w, h, d)
    dims[axis] = 2 * radius + 1
    w, h, d = dims
    a = np.ones((d, h, w), np.uint8)
    ends = [slice(None)] * 3
    for pos in (0, 2 * radius):
        ends[2 - axis] = pos
        a[tuple(ends)] = 0
    ref = orc.local_thickness_stages(a, w, h, d, 0)
    assert int(ref["squared_edt"].max()) == radius * radius
    got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got["propagated"], ref["propagated"])
    np.testing.assert_array_equal(got["thickness"], ref["thickness"])


@pytest.mark.parametrize("tiles", ["1", "2"])
@pytest.mark.parametrize("case", ["structure", "sparse_sites"])
def test_fill_packed_position_form_matches_oracle(case, tiles, monkeypatch):
    """The packed-position candidate test (taken above radius 64) forced at
    small radii, on the volumes the dot-product test normally serves: both
    forms give the oracle's map."""
    monkeypatch.setenv("BM_FILL_DOT", "0")
    monkeypatch.setenv("BM_FILL_TILES", tiles)
    if case == "structure":
        w, h, d = 40, 36, 44
        a = _structure(5, (d, h, w), 0.75)
    else:
        rng = np.random.default_rng(23)
        w, h, d = 72, 64, 56
        a = np.ones(w * h * d, np.uint8)
        a[rng.choice(a.size, 30, replace=False)] = 0
    ref = orc.local_thickness_stages(a, w, h, d, 0)
    got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got["propagated"], ref["propagated"])
    np.testing.assert_array_equal(got["thickness"], ref["thickness"])


@pytest.mark.parametrize("tiles", ["1", "2"])
@pytest.mark.parametrize("overrides", [{"BM_FILL_CHUNK_MAP": "0"}, {"BM_FILL_STAGE": "0"},
                                       {"BM_FILL_CHUNK_MAP": "0", "BM_FILL_STAGE": "0",
                                        "BM_FILL_CAP": "96"}],
                         ids=["binary-search chunks", "unstaged", "both, small rounds"])
def test_fill_slow_paths_match_oracle(overrides, tiles, monkeypatch):
    """The fill's paths for tiles with more list chunks than its descriptors
    hold (a binary search over the chunk offsets) and more candidates than its
    staging area holds (every round re-reads the lists), forced on an ordinary
    volume and a volume of a few sites: the oracle's map either way."""
    for k, v in overrides.items():
        monkeypatch.setenv(k, v)
    monkeypatch.setenv("BM_FILL_TILES", tiles)
    for case in ("structure", "sparse_sites"):
        if case == "structure":
            w, h, d = 40, 36, 44
            a = _structure(5, (d, h, w), 0.75)
        else:
            rng = np.random.default_rng(29)
            w, h, d = 72, 64, 56
            a = np.ones(w * h * d, np.uint8)
            a[rng.choice(a.size, 30, replace=False)] = 0
        ref = orc.local_thickness_stages(a, w, h, d, 0)
        got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
        np.testing.assert_array_equal(got["propagated"], ref["propagated"], err_msg=case)
        np.testing.assert_array_equal(got["thickness"], ref["thickness"], err_msg=case)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fill_axis_beyond_1024_tiles_matches_oracle(axis):
    """An axis of 8216 voxels (1027 tiles): tile coordinates past 10 bits in
    the fill's gathering of the tiles in reach."""
    dims = [6, 5, 7]
    dims[axis] = 8216
    w, h, d = dims
    rng = np.random.default_rng(40 + axis)
    a = (rng.random(w * h * d) > 0.04).astype(np.uint8)
    ref = orc.local_thickness_stages(a, w, h, d, 0)
    got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got["ridge_flags"], ref["ridge_flags"])
    np.testing.assert_array_equal(got["propagated"], ref["propagated"])
    np.testing.assert_array_equal(got["thickness"], ref["thickness"])


def test_analyze_large_radius_volume_matches_oracle():
    """A volume of the batched workload whose marrow distances exceed 511."""
    p = workloads.batched_params()[29]
    wl = workloads.make_volume(p)
    ref = orc.analyze_thickness_spacing(wl.bone, wl.roi, wl.width, wl.height, wl.depth,
                                        wl.spacing)
    got = bm.analyze_thickness_spacing(vol(wl.bone, wl.width, wl.height, wl.depth),
                                       vol(wl.roi, wl.width, wl.height, wl.depth), wl.spacing)
    _check_stats(got.thickness, ref.thickness, scale=wl.spacing[0])
    _check_stats(got.spacing, ref.spacing, scale=wl.spacing[0])


def test_analyze_large_workload_generator_matches_oracle():
    """The large workload's generator (sigma (5, 5, 9), BV/TV 0.12, ellipsoid
    ROI, 8 um) at 192x192x288: marrow radii up to ~24 voxels, so the fill runs
    with 9x9x9 tile neighbourhoods and mixed-radius buckets."""
    w, h, d = 192, 192, 288
    bone = synth.volume(w, h, d, (5.0, 5.0, 9.0), 0.12, 3)
    roi = synth.roi_ellipsoid(w, h, d, (w * 0.49, h * 0.47, d * 0.48))
    ref = orc.analyze_thickness_spacing(bone, roi, w, h, d, (8.0, 8.0, 8.0))
    got = bm.analyze_thickness_spacing(vol(bone, w, h, d), vol(roi, w, h, d), (8.0, 8.0, 8.0))
    assert ref.spacing.maximum / 8.0 > 40.0
    _check_stats(got.thickness, ref.thickness, scale=8.0)
    _check_stats(got.spacing, ref.spacing, scale=8.0)


def test_local_thickness_plain_small_matches_oracle():
    wl = workloads.plain_small()
    for phase in (bm.Phase.Foreground, bm.Phase.Background):
        got = bm.local_thickness(vol(wl.bone, wl.width, wl.height, wl.depth), phase)
        ref = orc.local_thickness(wl.bone, wl.width, wl.height, wl.depth, int(phase))
        np.testing.assert_array_equal(got, ref)


def test_local_thickness_is_deterministic():
    wl = workloads.plain_small()
    v = vol(wl.bone, wl.width, wl.height, wl.depth)
    a = bm.local_thickness(v, bm.Phase.Background)
    b = bm.local_thickness(v, bm.Phase.Background)
    np.testing.assert_array_equal(a, b)


# --- analyze (thickness / spacing part) ------------------------------------------
def _close(a, b):
    if isinstance(b, float) and math.isnan(b):
        return math.isnan(a)
    return abs(a - b) <= REL * abs(b)


def _check_stats(g, r, sequential=False, scale=None):
    check_stats(g, r, sequential, scale)


@pytest.mark.parametrize("order", [bm.StatsOrder.Exact, bm.StatsOrder.Sequential])
@pytest.mark.parametrize("mode", [bm.FeatureMode.Standard, bm.FeatureMode.Classic])
def test_analyze_plain_small_matches_oracle(mode, order):
    wl = workloads.plain_small()
    b = vol(wl.bone, wl.width, wl.height, wl.depth)
    r = vol(wl.roi, wl.width, wl.height, wl.depth)
    with bm.stats_order(order):
        got = bm.analyze_thickness_spacing(b, r, wl.spacing, mode)
    q = mode in (bm.FeatureMode.Standard, bm.FeatureMode.Experimental)
    ref = orc.analyze_thickness_spacing(wl.bone, wl.roi, wl.width, wl.height, wl.depth,
                                        wl.spacing, include_quantiles=q)
    seq = order == bm.StatsOrder.Sequential
    _check_stats(got.thickness, ref.thickness, seq, wl.spacing[0])
    _check_stats(got.spacing, ref.spacing, seq, wl.spacing[0])
    assert got.bone_volume == ref.bone_volume and got.total_volume == ref.total_volume
    assert got.bone_volume_fraction == ref.bone_volume_fraction


def test_analyze_cylinder_roi_matches_oracle():
    w, h, d = 96, 80, 72
    from bone_morphometry_amd import synth
    bone = synth.volume(w, h, d, 2.5, 0.2, 11)
    roi = synth.roi_cylinder(w, h, d, 70.0)
    ref = orc.analyze_thickness_spacing(bone, roi, w, h, d, (7.25, 7.25, 7.25))
    got = bm.analyze_thickness_spacing(vol(bone, w, h, d), vol(roi, w, h, d), (7.25, 7.25, 7.25))
    _check_stats(got.thickness, ref.thickness, scale=7.25)
    _check_stats(got.spacing, ref.spacing, scale=7.25)
    with bm.stats_order(bm.StatsOrder.Sequential):
        got = bm.analyze_thickness_spacing(vol(bone, w, h, d), vol(roi, w, h, d),
                                           (7.25, 7.25, 7.25))
    _check_stats(got.thickness, ref.thickness, True)
    _check_stats(got.spacing, ref.spacing, True)


def _ranks_sharing_top_bins(values):
    """Top-16-bit bins holding two of the quantile ranks' distinct 24-bit
    prefixes (the digit passes' shared-bin case)."""
    b = np.sort(np.asarray(values, np.float32).view(np.uint32))
    ranks = set()
    for q in (0.5, 0.1, 0.25, 0.75, 0.9):
        pos = q * (b.size - 1)
        ranks |= {int(np.floor(pos)), int(np.ceil(pos))}
    tops = {}
    for p in {int(b[r]) >> 8 for r in ranks}:
        tops.setdefault(p >> 8, set()).add(p)
    return sum(len(s) > 1 for s in tops.values())


def test_quantiles_with_shared_top_bins_match_oracle():
    """A marrow phase whose quantile ranks have two distinct 24-bit prefixes
    in one top-16-bit bin: the second digit pass classifies those values by
    searching the groups."""
    from bone_morphometry_amd import synth
    w, h, d = 48, 40, 36
    bone = synth.volume(w, h, d, 2.0, 0.3, 5)
    roi = synth.roi_cylinder(w, h, d, 36.0)
    t = orc.local_thickness(orc.phase_inside(bone, roi, 0), w, h, d, 0)
    assert _ranks_sharing_top_bins(t[t > 0]) >= 1
    ref = orc.analyze_thickness_spacing(bone, roi, w, h, d, (9.0, 9.0, 9.0))
    got = bm.analyze_thickness_spacing(vol(bone, w, h, d), vol(roi, w, h, d), (9.0, 9.0, 9.0))
    _check_stats(got.thickness, ref.thickness, scale=9.0)
    _check_stats(got.spacing, ref.spacing, scale=9.0)


def test_stats_order_rejects_unknown_values():
    with pytest.raises(bm.BmError, match="unknown statistics order 7"):
        bm.set_stats_order(7)
    assert bm.set_stats_order(bm.StatsOrder.Exact) == bm.StatsOrder.Exact


# --- radii beyond the gather: the scatter of the largest balls -------------------
@pytest.mark.parametrize("radius", [129, 160])
@pytest.mark.parametrize("axis", [0, 2])
def test_fill_scatter_beyond_gather_radius_matches_oracle(axis, radius):
    """Two planes of sites 2R apart: every ridge of the mid-plane has radius R
    above the gather's 128, so its ball is scattered; the voxels near the
    planes take the gathered smaller balls."""
    dims = [11, 10, 9]
    dims[axis] = 2 * radius + 1
    w, h, d = dims
    a = np.ones((d, h, w), np.uint8)
    ends = [slice(None)] * 3
    for pos in (0, 2 * radius):
        ends[2 - axis] = pos
        a[tuple(ends)] = 0
    ref = orc.local_thickness_stages(a, w, h, d, 0)
    assert int(ref["squared_edt"].max()) == radius * radius
    got = bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got["ridge_flags"], ref["ridge_flags"])
    np.testing.assert_array_equal(got["propagated"], ref["propagated"])
    np.testing.assert_array_equal(got["thickness"], ref["thickness"])
    assert got["fill_visits"] == ref["fill_visits"]


def test_fill_large_tube_cavity_matches_oracle():
    """A tube-shaped cavity of radius 130.5 across a 12-voxel slab (sites
    outside the tube only): the distance field does not change along z, so
    the ridges are the digital disk's medial set in every slice, with radii
    from 1 to above 128.  Balls up to radius 128 are gathered, the larger
    ones scattered afterwards; both meet in every voxel near the axis."""
    n, d = 264, 12
    c = (n - 1) / 2.0
    y, x = np.ogrid[0:n, 0:n]
    disk = ((x - c) ** 2 + (y - c) ** 2) <= 130.5 ** 2
    a = np.ascontiguousarray(np.broadcast_to(disk, (d, n, n)), np.uint8).reshape(-1)
    ref = orc.local_thickness_stages(a, n, n, d, 0)
    big = (ref["ridge_flags"] > 0) & (ref["squared_edt"] > 128 * 128)
    assert 0 < int(big.sum()) < int(ref["ridge_count"])
    got = bm.local_thickness_stages(vol(a, n, n, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got["ridge_flags"], ref["ridge_flags"])
    np.testing.assert_array_equal(got["propagated"], ref["propagated"])
    np.testing.assert_array_equal(got["thickness"], ref["thickness"])
    assert got["fill_visits"] == ref["fill_visits"]


@pytest.mark.parametrize("call", ["local_thickness", "stages", "analyze"])
def test_ridge_radius_beyond_list_range_raises(call, monkeypatch):
    """Ridges whose r^2 the tile lists cannot hold (2^23 and above, lowered
    here to 12 by the test override) fail every thickness path with
    BM_ERR_UNSUPPORTED instead of returning maps built from cut radii."""
    monkeypatch.setenv("BM_LIST_R2_LIMIT", "12")
    w, h, d = 40, 36, 44
    a = _structure(5, (d, h, w), 0.85)
    assert orc.edt_squared(a, w, h, d, 0).max() >= 12
    with pytest.raises(bm.BmError, match="limit for ridge radii") as e:
        if call == "local_thickness":
            bm.local_thickness(vol(a, w, h, d), bm.Phase.Foreground)
        elif call == "stages":
            bm.local_thickness_stages(vol(a, w, h, d), bm.Phase.Foreground)
        else:
            bm.analyze_thickness_spacing(vol(a, w, h, d), vol(np.ones_like(a), w, h, d),
                                         (1.0, 1.0, 1.0))
    assert e.value.code == 3
    monkeypatch.delenv("BM_LIST_R2_LIMIT")
    got = bm.local_thickness(vol(a, w, h, d), bm.Phase.Foreground)
    np.testing.assert_array_equal(got, orc.local_thickness(a, w, h, d, 0))


@pytest.mark.parametrize("bone,roi,mode,msg", [
    (np.zeros(8), np.zeros(8), bm.FeatureMode.Standard, "ROI mask contains no included voxels"),
    (np.zeros(8), np.ones(8), bm.FeatureMode.Standard, "ROI contains no bone voxels"),
    (np.ones(8), np.ones(8), bm.FeatureMode.Standard, "ROI contains no marrow voxels"),
    (np.eye(2).reshape(-1).repeat(2), np.ones(8), bm.FeatureMode.Bonej,
     "BoneJ compatibility mode requires an ROI mask with both inside and outside voxels"),
])
def test_analyze_error_precedence(bone, roi, mode, msg):
    with pytest.raises(bm.BmError, match=msg):
        bm.analyze_thickness_spacing(vol(bone, 2, 2, 2), vol(roi, 2, 2, 2), (1, 1, 1), mode)
