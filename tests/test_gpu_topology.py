"""GPU parity of the rest of analyze (src/topology.rs, src/metrics.rs:134-202)
through the C ABI: cubical measures, Betti numbers, BoneJ connectivity,
interface faces, the ROI-boundary fraction, and the whole Metrics of analyze.
The counts are integers and the f64 fields follow from them by the
reference's formulas, so all must equal the oracle's exactly."""
import numpy as np
import pytest
from scipy import ndimage

import _oracle as orc
from _stats import check_stats
import bone_morphometry_amd as bm
from bone_morphometry_amd import synth, workloads

pytestmark = pytest.mark.gpu


def vol(a):
    a = np.asarray(a, np.uint8)
    d, h, w = a.shape
    return bm.BinaryVolume(a.reshape(-1), w, h, d)


def points(pts, n=5):
    v = np.zeros((n, n, n), np.uint8)
    for (x, y, z) in pts:
        v[z, y, x] = 1
    return v


# --- the reference's unit tests ---------------------------------------------------------
def test_single_and_two_voxel_known_answers():
    full = vol(np.ones((5, 5, 5)))
    c = bm.cubical_measures(vol(points([(2, 2, 2)])))
    assert c.raw_euler == 1 and abs(c.mean_breadth_lattice_units - 1.5) < 1e-12
    assert bm.bone_marrow_interface_faces(vol(points([(2, 2, 2)])), full) == 6
    c = bm.cubical_measures(vol(points([(1, 1, 1), (2, 1, 1)])))
    assert c.raw_euler == 1 and abs(c.mean_breadth_lattice_units - 2.0) < 1e-12
    assert bm.bone_marrow_interface_faces(vol(points([(1, 1, 1), (2, 1, 1)])), full) == 10


def test_bonej_known_answers():
    n = 9
    v = np.zeros((n, n, n), np.uint8)
    for z in range(2, 7):
        for y in range(2, 7):
            for x in range(2, 7):
                if (x in (2, 6)) + (y in (2, 6)) + (z in (2, 6)) >= 2:
                    v[z, y, x] = 1
    raw = bm.cubical_measures(vol(v)).raw_euler
    b = bm.bonej_connectivity(vol(v), raw)
    assert raw == -4 and (b.edge_correction, b.delta_chi, b.connectivity) == (0.0, -4.0, 5.0)
    corner = vol(points([(0, 0, 0)]))
    raw = bm.cubical_measures(corner).raw_euler
    b = bm.bonej_connectivity(corner, raw)
    assert raw == 1 and (b.edge_correction, b.delta_chi, b.connectivity) == (0.875, 0.125, 0.875)


def test_topology_known_answers():
    for pts, want in (([(1, 1, 1), (2, 2, 2)], (1, 1, 0, 0)), ([(1, 1, 1), (3, 3, 3)], (2, 2, 0, 0))):
        v = vol(points(pts))
        t = bm.analyze_topology(v, bm.cubical_measures(v).raw_euler)
        assert (t.raw_euler, t.beta0, t.beta1, t.beta2) == want
    h = np.zeros((5, 5, 5), np.uint8)
    h[1:4, 1:4, 1:4] = 1
    h[2, 2, 2] = 0
    t = bm.analyze_topology(vol(h), bm.cubical_measures(vol(h)).raw_euler)
    assert (t.raw_euler, t.beta0, t.beta1, t.beta2) == (2, 1, 0, 1)


def test_interface_cut_face_and_metrics_known_answers():
    roi = np.zeros((5, 5, 5), np.uint8)
    roi[1:4, 1:4, 1:4] = 1
    bone = points([(1, 2, 2)])
    assert bm.bone_marrow_interface_faces(vol(bone), vol(roi)) == 5
    m = bm.analyze(vol(bone), vol(roi), (1.0, 1.0, 1.0), bm.FeatureMode.Refined)
    assert m.bone_volume == 1.0 and m.total_volume == 27.0
    assert abs(m.bone_volume_fraction - 1.0 / 27.0) < 1e-12
    assert m.geometry.bone_surface == 5.0
    t = m.topology
    assert (t.raw_euler, t.beta0, t.beta1, t.beta2) == (1, 1, 0, 0)
    assert m.roi_boundary_bone_fraction == 1.0


def test_errors():
    with pytest.raises(bm.BmError, match=r"^computed a negative beta1 \(-99\); check the 26/6 "
                                          r"topology convention$"):
        bm.analyze_topology(vol(points([(2, 2, 2)])), 100)
    with pytest.raises(bm.BmError, match="^bone dimensions 5x5x5 do not match ROI dimensions "
                                          "5x5x6$"):
        bm.bone_marrow_interface_faces(vol(points([(2, 2, 2)])), vol(np.ones((6, 5, 5))))
    roi = np.zeros((5, 5, 5), np.uint8)
    with pytest.raises(bm.BmError, match="^ROI mask contains no included voxels$"):
        bm.analyze(vol(points([(2, 2, 2)])), vol(roi), (1.0, 1.0, 1.0))


# --- random volumes against the oracle ------------------------------------------------------
def random_volume(seed, shape, frac, smooth):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape)
    if smooth:
        f = ndimage.uniform_filter(f, smooth)
    return (f > np.quantile(f, 1 - frac)).astype(np.uint8)


CASES = [(0, (7, 9, 11), 0.3, 0), (1, (12, 10, 14), 0.5, 3), (2, (2, 2, 2), 0.5, 0),
         (3, (2, 3, 17), 0.6, 0), (4, (33, 65, 70), 0.45, 3), (5, (20, 5, 9), 0.7, 3),
         (6, (64, 64, 64), 0.2, 5), (7, (40, 130, 66), 0.85, 3), (8, (50, 50, 50), 0.05, 0)]


@pytest.mark.parametrize("seed,shape,frac,smooth", CASES)
def test_volume_measures_match_oracle(seed, shape, frac, smooth):
    v = random_volume(seed, shape, frac, smooth)
    d, h, w = shape
    c = bm.cubical_measures(vol(v))
    oc = orc.cubical_measures(v, w, h, d)
    assert (c.raw_euler, c.mean_breadth_lattice_units) == (oc.raw_euler,
                                                           oc.mean_breadth_lattice_units)
    t = bm.analyze_topology(vol(v), c.raw_euler)
    ot = orc.analyze_topology(v, w, h, d, c.raw_euler)
    assert (t.beta0, t.beta1, t.beta2) == (ot.beta0, ot.beta1, ot.beta2)
    b = bm.bonej_connectivity(vol(v), c.raw_euler)
    ob = orc.bonej_connectivity(v, w, h, d, c.raw_euler)
    assert (b.edge_correction, b.delta_chi, b.connectivity) == (ob.edge_correction, ob.delta_chi,
                                                                ob.connectivity)
    roi = random_volume(seed + 100, shape, 0.8, 2)
    assert bm.bone_marrow_interface_faces(vol(v), vol(roi)) == orc.interface_faces(v, roi, w, h, d)


def test_labels_are_deterministic_on_many_components():
    # salt-and-pepper: ~10^5 components of each phase, unions landing in any order
    v = random_volume(9, (60, 70, 80), 0.3, 0)
    c = bm.cubical_measures(vol(v))
    first = bm.analyze_topology(vol(v), c.raw_euler)
    for _ in range(2):
        assert bm.analyze_topology(vol(v), c.raw_euler) == first
    ot = orc.analyze_topology(v, 80, 70, 60, c.raw_euler)
    assert (first.beta0, first.beta2) == (ot.beta0, ot.beta2)


def check_metrics(m, bone, roi, w, h, d, spacing, mode):
    rest = orc.analyze_rest(bone, roi, w, h, d, spacing)
    t = m.topology
    assert (t.raw_euler, t.beta0, t.beta1, t.beta2) == (rest.topology.raw_euler, rest.topology.beta0,
                                                        rest.topology.beta1, rest.topology.beta2)
    b = m.bonej_connectivity
    assert (b.edge_correction, b.delta_chi, b.connectivity) == (
        rest.bonej.edge_correction, rest.bonej.delta_chi, rest.bonej.connectivity)
    for f in ("bone_volume", "total_volume", "bone_volume_fraction", "bonej_connectivity_density",
              "connectivity", "connectivity_density", "roi_boundary_bone_fraction"):
        assert getattr(m, f) == getattr(rest, f), f
    g = m.geometry
    for f in ("bone_surface", "bone_surface_to_bone_volume", "bone_surface_to_total_volume",
              "mean_breadth"):
        assert getattr(g, f) == getattr(rest, f), f
    ts = orc.analyze_thickness_spacing(bone, roi, w, h, d, spacing,
                                       include_quantiles=(mode == bm.FeatureMode.Standard))
    for got, ref in ((m.thickness, ts.thickness), (m.spacing, ts.spacing)):
        check_stats(got, ref, scale=spacing[0])
    return ts


def test_analyze_small_workload_matches_oracle():
    wl = workloads.plain_small()
    m = bm.analyze(vol(wl.bone.reshape(wl.depth, wl.height, wl.width)),
                   vol(wl.roi.reshape(wl.depth, wl.height, wl.width)), wl.spacing)
    check_metrics(m, wl.bone, wl.roi, wl.width, wl.height, wl.depth, wl.spacing,
                  bm.FeatureMode.Standard)


def test_analyze_cylinder_roi_matches_oracle():
    w, h, d = 96, 80, 70
    bone = synth.volume(w, h, d, 2.5, 0.2, 11)
    roi = synth.roi_cylinder(w, h, d, 70.0)
    m = bm.analyze(bm.BinaryVolume(bone, w, h, d), bm.BinaryVolume(roi, w, h, d),
                   (12.0, 12.0, 12.0), bm.FeatureMode.Classic)
    check_metrics(m, bone, roi, w, h, d, (12.0, 12.0, 12.0), bm.FeatureMode.Classic)


def test_analyze_512_workload_matches_oracle():
    wl = workloads.plain_512()
    from bone_morphometry_amd import hip
    d_bone, d_roi = hip.DeviceBuffer(wl.n), hip.DeviceBuffer(wl.n)
    d_bone.upload(wl.bone)
    d_roi.upload(wl.roi)
    m = bm.analyze_device(d_bone.value, d_roi.value, wl.width, wl.height, wl.depth, wl.spacing)
    ts = check_metrics(m, wl.bone, wl.roi, wl.width, wl.height, wl.depth, wl.spacing,
                       bm.FeatureMode.Standard)
    # the reference's summation order: mean, SD and CV equal to the oracle's
    with bm.stats_order(bm.StatsOrder.Sequential):
        m = bm.analyze_device(d_bone.value, d_roi.value, wl.width, wl.height, wl.depth,
                              wl.spacing)
    check_stats(m.thickness, ts.thickness, sequential=True)
    check_stats(m.spacing, ts.spacing, sequential=True)
