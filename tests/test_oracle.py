"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is independent of the oracle's code: the reference's own unit tests
restated (known answers), scipy-derived fixtures, brute force on tiny volumes,
closed forms, and invariants that a plausible mistake (a dropped term, a wrong
index, a swapped offset) breaks.
"""
import itertools
import math
import os

import numpy as np
import pytest

import _oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "edt_scipy.npz")


def brute_edt(sel):
    """min over non-selected voxels of the squared distance (d, h, w array)."""
    sites = np.argwhere(~sel)
    out = np.zeros(sel.shape, np.int64)
    for p in np.argwhere(sel):
        out[tuple(p)] = ((sites - p) ** 2).sum(axis=1).min()
    return out


def isqrt(v):
    return math.isqrt(v)


def brute_template(r2, dx, dy, dz):
    """max of (i+dx)^2+(j+dy)^2+(k+dz)^2 over lattice points i,j,k >= 0 in the ball."""
    r = isqrt(r2)
    best = 0
    for i in range(r + 1):
        for j in range(r + 1):
            for k in range(r + 1):
                if i * i + j * j + k * k <= r2:
                    best = max(best, (i + dx) ** 2 + (j + dy) ** 2 + (k + dz) ** 2)
    return best


# --- src/volume.rs:123-136 ---------------------------------------------------
def test_phase_inside_distinguishes_marrow_from_outside_roi():
    bone = np.array([0, 1, 0, 1, 0, 1, 0, 1], np.uint8)
    roi = np.array([0, 0, 1, 1, 1, 1, 0, 0], np.uint8)
    assert orc.phase_inside(bone, roi, True).tolist() == [0, 0, 0, 1, 0, 1, 0, 0]
    assert orc.phase_inside(bone, roi, False).tolist() == [0, 0, 1, 0, 1, 0, 0, 0]


# --- src/edt.rs:240-274 -------------------------------------------------------
def test_edt_center_background():
    data = np.ones(27, np.uint8)
    data[13] = 0
    edt = orc.edt_squared(data, 3, 3, 3, 0)
    assert edt[13] == 0 and edt[12] == 1 and edt[0] == 3


def test_background_phase_uses_foreground_as_sites():
    data = np.zeros(27, np.uint8)
    data[13] = 1
    edt = orc.edt_squared(data, 3, 3, 3, 1)
    assert edt[13] == 0 and edt[12] == 1 and edt[0] == 3


def test_edt_all_selected_is_an_error():
    with pytest.raises(orc.OracleError, match="undefined for an all-foreground volume"):
        orc.edt_squared(np.ones(8, np.uint8), 2, 2, 2, 0)
    with pytest.raises(orc.OracleError, match="undefined for an all-background volume"):
        orc.edt_squared(np.zeros(8, np.uint8), 2, 2, 2, 1)


def test_edt_rejects_thin_volumes():
    with pytest.raises(orc.OracleError, match="at least 2"):
        orc.edt_squared(np.ones(4, np.uint8), 1, 2, 2, 0)


def test_edt_matches_scipy_fixtures():
    g = np.load(GOLDEN)
    k = 0
    while f"vol_{k}" in g:
        w, h, d = (int(v) for v in g[f"shape_{k}"])
        got = orc.edt_squared(g[f"vol_{k}"], w, h, d, 0)
        np.testing.assert_array_equal(got, g[f"d2_{k}"], err_msg=f"case {k} {w}x{h}x{d}")
        k += 1
    assert k >= 10


@pytest.mark.parametrize("seed", range(6))
def test_edt_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    w, h, d = rng.integers(2, 9, size=3)
    sel = rng.random((d, h, w)) < rng.uniform(0.3, 0.95)
    if sel.all():
        sel[0, 0, 0] = False
    got = orc.edt_squared(sel.astype(np.uint8).reshape(-1), w, h, d, 0)
    np.testing.assert_array_equal(got.reshape(d, h, w), brute_edt(sel))


def test_edt_single_site_closed_form():
    w, h, d = 13, 7, 9
    sel = np.ones((d, h, w), bool)
    p = (4, 2, 9)  # z, y, x
    sel[p] = False
    zz, yy, xx = np.indices(sel.shape)
    expect = (zz - p[0]) ** 2 + (yy - p[1]) ** 2 + (xx - p[2]) ** 2
    got = orc.edt_squared(sel.astype(np.uint8).reshape(-1), w, h, d, 0).reshape(d, h, w)
    np.testing.assert_array_equal(got, expect)


# --- src/local_thickness.rs:395-401 and the brute-force lattice maximum -------
def test_search_template_matches_known_single_voxel_radius():
    t = orc.templates([1])
    assert t[:, 0].tolist() == [4, 5, 6]


def test_templates_match_lattice_brute_force():
    values = list(range(0, 90)) + [97, 120, 169, 200, 386]
    t = orc.templates(values)
    for i, r2 in enumerate(values):
        assert t[0, i] == brute_template(r2, 1, 0, 0), r2
        assert t[1, i] == brute_template(r2, 1, 1, 0), r2
        assert t[2, i] == brute_template(r2, 1, 1, 1), r2


# --- src/local_thickness.rs:364-393 -------------------------------------------
def test_isolated_voxel_has_diameter_two():
    data = np.zeros(27, np.uint8)
    data[13] = 1
    th = orc.local_thickness(data, 3, 3, 3, 0)
    assert abs(th[13] - 2.0) < 1e-6
    assert (th > 0).sum() == 1


def test_background_phase_matches_complemented_foreground_phase():
    data = np.zeros((5, 5, 5), np.uint8)
    data[1:4, 1:4, 1:4] = 1
    direct = orc.local_thickness(data.reshape(-1), 5, 5, 5, 1)
    comp = orc.local_thickness((data == 0).astype(np.uint8).reshape(-1), 5, 5, 5, 0)
    np.testing.assert_array_equal(direct, comp)


# --- invariants of the local-thickness construction ---------------------------
NB26 = [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0), (0, 1, -1),
        (0, 1, 1), (1, -1, 0), (1, 1, 0), (-1, 0, 1), (1, 0, 1), (0, -1, -1), (0, -1, 1),
        (-1, -1, 0), (-1, 1, 0), (-1, 0, -1), (1, 0, -1), (1, 1, 1), (1, -1, 1), (-1, 1, 1),
        (-1, -1, 1), (1, 1, -1), (1, -1, -1), (-1, 1, -1), (-1, -1, -1)]


def _random_structure(seed, shape, frac):
    """Smooth random binary structure (d, h, w) with a given selected fraction."""
    import scipy.ndimage as ndi
    rng = np.random.default_rng(seed)
    f = ndi.gaussian_filter(rng.standard_normal(shape), 1.5)
    return f > np.quantile(f, 1.0 - frac)


@pytest.mark.parametrize("seed,frac", [(0, 0.3), (1, 0.6), (2, 0.85)])
def test_propagation_from_ridges_equals_propagation_from_all_voxels(seed, frac):
    """Every non-ridge ball lies inside a neighbour's larger ball, so the ridge
    set is sufficient: a filter that drops a needed centre changes the map."""
    sel = _random_structure(seed, (18, 20, 22), frac)
    d, h, w = sel.shape
    v = sel.astype(np.uint8).reshape(-1)
    a = orc.local_thickness_stages(v, w, h, d, 0)
    b = orc.local_thickness_stages(v, w, h, d, 0, propagate_from_all=True)
    assert a["ridge_count"] < b["ridge_count"]
    np.testing.assert_array_equal(a["propagated"], b["propagated"])
    np.testing.assert_array_equal(a["thickness"], b["thickness"])


@pytest.mark.parametrize("seed,frac", [(3, 0.4), (4, 0.7)])
def test_ridge_set_is_the_template_rule(seed, frac):
    """Ridge flags equal the rule of src/local_thickness.rs:131-162 evaluated
    with brute-force templates, and contain every local maximum of d2."""
    sel = _random_structure(seed, (12, 13, 14), frac)
    d, h, w = sel.shape
    st = orc.local_thickness_stages(sel.astype(np.uint8).reshape(-1), w, h, d, 0)
    d2 = st["squared_edt"].reshape(d, h, w).astype(np.int64)
    flags = st["ridge_flags"].reshape(d, h, w)
    tcache = {}
    for z, y, x in itertools.product(range(d), range(h), range(w)):
        r2 = int(d2[z, y, x])
        if r2 == 0:
            assert flags[z, y, x] == 0
            continue
        if r2 not in tcache:
            tcache[r2] = [brute_template(r2, 1, 0, 0), brute_template(r2, 1, 1, 0),
                          brute_template(r2, 1, 1, 1)]
        ridge, local_max = True, True
        for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
            if (dz, dy, dx) == (0, 0, 0):
                continue
            nz, ny, nx = z + dz, y + dy, x + dx
            if not (0 <= nz < d and 0 <= ny < h and 0 <= nx < w):
                continue
            c = (dx != 0) + (dy != 0) + (dz != 0)
            if d2[nz, ny, nx] >= tcache[r2][c - 1]:
                ridge = False
            if d2[nz, ny, nx] > r2:
                local_max = False
        assert flags[z, y, x] == int(ridge), (z, y, x)
        if local_max:
            assert ridge


def test_fill_visits_of_single_voxels():
    """W_fill of one selected voxel: its ball r^2 = 1 is the voxel and its six
    face neighbours, cut by the array at a corner (src/local_thickness.rs:182-201)."""
    for pos, want in (((2, 2, 2), 7), ((0, 0, 0), 4), ((0, 2, 4), 5)):
        v = np.zeros((5, 5, 5), np.uint8)
        v[pos] = 1
        st = orc.local_thickness_stages(v.reshape(-1), 5, 5, 5, 0)
        assert st["ridge_count"] == 1 and st["fill_visits"] == want, pos


@pytest.mark.parametrize("seed,frac", [(7, 0.3), (8, 0.75)])
def test_fill_visits_equal_brute_force_ball_counts(seed, frac):
    """W_fill = sum over ridge voxels of the lattice points within r^2 = d2 of
    the centre inside the array, counted here over the whole grid."""
    sel = _random_structure(seed, (11, 12, 13), frac)
    d, h, w = sel.shape
    st = orc.local_thickness_stages(sel.astype(np.uint8).reshape(-1), w, h, d, 0)
    d2 = st["squared_edt"].reshape(d, h, w).astype(np.int64)
    ridges = np.argwhere(st["ridge_flags"].reshape(d, h, w) != 0)
    Z, Y, X = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    total = 0
    for z, y, x in ridges:
        total += int(((Z - z) ** 2 + (Y - y) ** 2 + (X - x) ** 2 <= d2[z, y, x]).sum())
    assert st["ridge_count"] == len(ridges) > 0
    assert st["fill_visits"] == total


@pytest.mark.parametrize("seed,frac", [(5, 0.35), (6, 0.8)])
def test_cleanup_is_order_independent_and_bounded(seed, frac):
    sel = _random_structure(seed, (14, 15, 16), frac)
    d, h, w = sel.shape
    v = sel.astype(np.uint8).reshape(-1)
    a = orc.local_thickness_stages(v, w, h, d, 0)
    b = orc.local_thickness_stages(v, w, h, d, 0, reverse_cleanup=True)
    np.testing.assert_array_equal(a["thickness"], b["thickness"])
    th = a["thickness"].reshape(d, h, w)
    m = a["propagated"].reshape(d, h, w)
    diam = (2.0 * np.sqrt(m.astype(np.float64))).astype(np.float32)
    assert (th[~sel] == 0).all()
    assert (th[sel] >= 2.0).all()
    # independent evaluation of the clean-up rule (src/local_thickness.rs:214-277)
    pad = np.pad(m, 1, constant_values=1)  # out of bounds is never background
    zero_nb = np.zeros(m.shape, bool)
    for dx, dy, dz in NB26:  # the reference's tuples are (dx, dy, dz)
        zero_nb |= pad[1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] == 0
    surface = (m != 0) & zero_nb
    for z, y, x in zip(*np.nonzero(sel)):
        if not surface[z, y, x]:
            assert th[z, y, x] == diam[z, y, x]
            continue
        s, cnt, lo, hi = np.float32(0.0), 0, np.inf, -np.inf
        for dx, dy, dz in NB26:
            nz, ny, nx = z + dz, y + dy, x + dx
            if 0 <= nz < d and 0 <= ny < h and 0 <= nx < w:
                if m[nz, ny, nx] != 0 and not surface[nz, ny, nx]:
                    s = np.float32(s + diam[nz, ny, nx])
                    cnt += 1
                    lo, hi = min(lo, diam[nz, ny, nx]), max(hi, diam[nz, ny, nx])
        expect = np.float32(s / np.float32(cnt)) if cnt else diam[z, y, x]
        assert th[z, y, x] == expect, (z, y, x)
        if cnt:  # a mean of its neighbours, up to the f32 rounding of the sum
            tol = 26 * np.finfo(np.float32).eps
            assert lo * (1 - tol) <= th[z, y, x] <= hi * (1 + tol)


# --- src/metrics.rs:304-326 ---------------------------------------------------
def test_quantiles_use_numpy_linear_interpolation():
    v = [1.0, 2.0, 3.0, 4.0]
    assert abs(orc.quantile(v, 0.1, 1.0) - 1.3) < 1e-6
    assert abs(orc.quantile(v, 0.5, 1.0) - 2.5) < 1e-6
    assert abs(orc.quantile(v, 0.9, 1.0) - 3.7) < 1e-6
    r = np.random.default_rng(0).random(101).astype(np.float32) + 0.5
    for q in (0.1, 0.25, 0.5, 0.75, 0.9):
        assert abs(orc.quantile(r, q, 1.0) - np.quantile(r.astype(np.float64), q)) < 1e-6


def test_distribution_statistics_include_cv():
    st = orc.nonzero_stats([0.0, 1.0, 2.0, 3.0, 4.0], 2.0, True)
    assert abs(st.mean - 5.0) < 1e-12
    assert abs(st.median - 5.0) < 1e-12
    assert abs(st.percentile_10 - 2.6) < 1e-6
    assert abs(st.percentile_90 - 7.4) < 1e-6
    assert abs(st.maximum - 8.0) < 1e-12
    assert abs(st.standard_deviation - 2.581988897471611) < 1e-12
    assert abs(st.coefficient_of_variation - st.standard_deviation / 5.0) < 1e-12


def test_stats_closed_form_on_random_values():
    rng = np.random.default_rng(7)
    v = (rng.random(20000) * 30 + 2).astype(np.float32)
    v[::7] = 0.0
    v[3] = np.inf  # non-finite values are skipped
    st = orc.nonzero_stats(v, 10.5, False)
    x = (v[np.isfinite(v) & (v > 0)] * np.float32(10.5)).astype(np.float64)
    assert st.count == x.size
    assert abs(st.mean - x.mean()) <= 1e-12 * x.mean()
    assert abs(st.standard_deviation - x.std(ddof=1)) <= 1e-9 * x.std(ddof=1)
    assert st.maximum == x.max()
    assert math.isnan(st.median)


# --- src/metrics.rs:64-132 ----------------------------------------------------
def test_roi_controls_volume():
    bone = np.zeros(125, np.uint8)
    bone[(2 * 5 + 2) * 5 + 1] = 1
    roi = np.zeros((5, 5, 5), np.uint8)
    roi[1:4, 1:4, 1:4] = 1
    m = orc.analyze_thickness_spacing(bone, roi.reshape(-1), 5, 5, 5, (1.0, 1.0, 1.0))
    assert m.bone_volume == 1.0 and m.total_volume == 27.0
    assert abs(m.bone_volume_fraction - 1.0 / 27.0) < 1e-12


@pytest.mark.parametrize("bone,roi,spacing,bonej,msg", [
    (np.zeros(8), np.zeros(8), (1, 1, 1), False, "ROI mask contains no included voxels"),
    (np.zeros(8), np.ones(8), (1, 1, 1), False, "ROI contains no bone voxels"),
    (np.ones(8), np.ones(8), (1, 1, 1), False, "ROI contains no marrow voxels"),
    (np.eye(2).reshape(-1).repeat(2), np.ones(8), (1, 1, 1), True,
     "BoneJ compatibility mode requires an ROI mask with both inside and outside voxels"),
    (np.ones(8), np.ones(8), (1, 0, 1), False, "voxel spacings must be finite and positive"),
])
def test_analyze_errors(bone, roi, spacing, bonej, msg):
    with pytest.raises(orc.OracleError, match=msg):
        orc.analyze_thickness_spacing(bone.astype(np.uint8), roi.astype(np.uint8), 2, 2, 2,
                                      spacing, bonej_mode=bonej)
