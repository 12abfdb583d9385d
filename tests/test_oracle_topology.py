"""Pins the oracle's restatement of the rest of analyze (src/topology.rs,
src/metrics.rs:134-202): cubical cell counts, Betti numbers, BoneJ's stack
edge correction, interface faces and the ROI-boundary fraction.

Independent checks: the reference's own unit tests (topology.rs:703-824,
metrics.rs:329-355), scipy.ndimage.label for the Betti numbers, and the
counts rebuilt with whole-array numpy operations on a zero-padded copy (a
different formulation from the oracle's per-voxel loops).
"""
import numpy as np
import pytest
from scipy import ndimage

import _oracle as orc


def points(pts, n=5):
    v = np.zeros((n, n, n), np.uint8)
    for (x, y, z) in pts:
        v[z, y, x] = 1
    return v


def full(n=5):
    return np.ones((n, n, n), np.uint8)


# --- the reference's unit tests -----------------------------------------------------
def test_single_voxel_known_answer():
    v = points([(2, 2, 2)])
    c = orc.cubical_measures(v, 5, 5, 5)
    assert c.raw_euler == 1 and abs(c.mean_breadth_lattice_units - 1.5) < 1e-12
    assert orc.interface_faces(v, full(), 5, 5, 5) == 6


def test_two_voxels_known_answer():
    v = points([(1, 1, 1), (2, 1, 1)])
    c = orc.cubical_measures(v, 5, 5, 5)
    assert c.raw_euler == 1 and abs(c.mean_breadth_lattice_units - 2.0) < 1e-12
    assert orc.interface_faces(v, full(), 5, 5, 5) == 10


def test_bonej_box_frame_known_answer():
    n = 9
    v = np.zeros((n, n, n), np.uint8)
    for z in range(2, 7):
        for y in range(2, 7):
            for x in range(2, 7):
                if (x in (2, 6)) + (y in (2, 6)) + (z in (2, 6)) >= 2:
                    v[z, y, x] = 1
    raw = orc.cubical_measures(v, n, n, n).raw_euler
    b = orc.bonej_connectivity(v, n, n, n, raw)
    assert raw == -4
    assert (b.edge_correction, b.delta_chi, b.connectivity) == (0.0, -4.0, 5.0)


def test_bonej_corner_voxel_known_answer():
    v = points([(0, 0, 0)])
    raw = orc.cubical_measures(v, 5, 5, 5).raw_euler
    b = orc.bonej_connectivity(v, 5, 5, 5, raw)
    assert raw == 1
    assert (b.edge_correction, b.delta_chi, b.connectivity) == (0.875, 0.125, 0.875)


def test_interface_excludes_roi_cut_face_known_answer():
    roi = np.zeros((5, 5, 5), np.uint8)
    roi[1:4, 1:4, 1:4] = 1
    assert orc.interface_faces(points([(1, 2, 2)]), roi, 5, 5, 5) == 5


@pytest.mark.parametrize("pts,raw,b0,b1,b2", [
    ([(1, 1, 1), (2, 2, 2)], 1, 1, 0, 0),   # corner-touching: one 26-component
    ([(1, 1, 1), (3, 3, 3)], 2, 2, 0, 0),   # separated
])
def test_components_known_answers(pts, raw, b0, b1, b2):
    v = points(pts)
    c = orc.cubical_measures(v, 5, 5, 5)
    t = orc.analyze_topology(v, 5, 5, 5, c.raw_euler)
    assert (c.raw_euler, t.beta0, t.beta1, t.beta2) == (raw, b0, b1, b2)


def test_hollow_cube_known_answer():
    v = np.zeros((5, 5, 5), np.uint8)
    v[1:4, 1:4, 1:4] = 1
    v[2, 2, 2] = 0
    c = orc.cubical_measures(v, 5, 5, 5)
    t = orc.analyze_topology(v, 5, 5, 5, c.raw_euler)
    assert (c.raw_euler, t.beta0, t.beta1, t.beta2) == (2, 1, 0, 1)


def test_analyze_rest_known_answer():
    # metrics.rs:329-355
    bone = points([(1, 2, 2)])
    roi = np.zeros((5, 5, 5), np.uint8)
    roi[1:4, 1:4, 1:4] = 1
    m = orc.analyze_rest(bone, roi, 5, 5, 5, (1.0, 1.0, 1.0))
    assert m.bone_volume == 1.0 and m.total_volume == 27.0
    assert abs(m.bone_volume_fraction - 1.0 / 27.0) < 1e-12
    assert m.bone_surface == 5.0
    t = m.topology
    assert (t.raw_euler, t.beta0, t.beta1, t.beta2) == (1, 1, 0, 0)
    assert m.roi_boundary_bone_fraction == 1.0


# --- independent whole-array formulations -------------------------------------------
def padded(v):
    d, h, w = v.shape
    p = np.zeros((d + 2, h + 2, w + 2), bool)
    p[1:-1, 1:-1, 1:-1] = v != 0
    return p


def np_cubical(v):
    d, h, w = v.shape
    b = v != 0
    p = padded(v)
    cubes = int(b.sum())
    shared = int((b[:, :, 1:] & b[:, :, :-1]).sum() + (b[:, 1:, :] & b[:, :-1, :]).sum()
                 + (b[1:] & b[:-1]).sum())
    faces = 6 * cubes - shared
    vert = np.zeros((d + 1, h + 1, w + 1), bool)
    ex = np.zeros((d + 1, h + 1, w), bool)
    ey = np.zeros((d + 1, h, w + 1), bool)
    ez = np.zeros((d, h + 1, w + 1), bool)
    for i in (0, 1):
        for j in (0, 1):
            ex |= p[i:i + d + 1, j:j + h + 1, 1:w + 1]
            ey |= p[i:i + d + 1, 1:h + 1, j:j + w + 1]
            ez |= p[1:d + 1, i:i + h + 1, j:j + w + 1]
            for k in (0, 1):
                vert |= p[i:i + d + 1, j:j + h + 1, k:k + w + 1]
    edges = int(ex.sum() + ey.sum() + ez.sum())
    return int(vert.sum()), edges, faces, cubes


def np_bonej_counts(v):
    """stack_vertices, stack_edges, stack_faces, face_vertices, face_edges, edge_vertices."""
    d, h, w = v.shape
    b = (v != 0).astype(np.int64)
    p = padded(v)
    Z, Y, X = (0, d - 1), (0, h - 1), (0, w - 1)
    sv = sum(b[z, y, x] for z in Z for y in Y for x in X)
    se = (sum(b[z, y, 1:w - 1].sum() for z in Z for y in Y)
          + sum(b[z, 1:h - 1, x].sum() for z in Z for x in X)
          + sum(b[1:d - 1, y, x].sum() for y in Y for x in X))
    sf = (sum(b[z, 1:h - 1, 1:w - 1].sum() for z in Z) + sum(b[1:d - 1, y, 1:w - 1].sum() for y in Y)
          + sum(b[1:d - 1, 1:h - 1, x].sum() for x in X))
    fv = 0
    for z in Z:
        pl = p[z + 1]
        lat = np.zeros((h + 1, w + 1), bool)
        for i in (0, 1):
            for j in (0, 1):
                lat |= pl[i:i + h + 1, j:j + w + 1]
        fv += lat.sum()
    for x in X:
        pl = p[:, :, x + 1]
        lat = np.zeros((d + 1, h + 1), bool)
        for i in (0, 1):
            for j in (0, 1):
                lat |= pl[i:i + d + 1, j:j + h + 1]
        fv += lat[1:d].sum()
    for y in Y:
        pl = p[:, y + 1, :]
        lat = np.zeros((d + 1, w + 1), bool)
        for i in (0, 1):
            for j in (0, 1):
                lat |= pl[i:i + d + 1, j:j + w + 1]
        fv += lat[1:d, 1:w].sum()
    fe = 0
    for z in Z:
        pl = p[z + 1]
        c = pl[1:h + 2, 1:w + 2]
        fe += np.where(c, 2, pl[0:h + 1, 1:w + 2].astype(int) + pl[1:h + 2, 0:w + 1]).sum()
    for y in Y:
        pl = p[:, y + 1, :]
        fe += (b[1:, y, :] | b[:-1, y, :]).sum() + (pl[1:d + 1, 1:w + 2] | pl[1:d + 1, 0:w + 1]).sum()
    for x in X:
        fe += (b[1:, :, x] | b[:-1, :, x]).sum() + (b[:, 1:, x] | b[:, :-1, x]).sum()
    ev = (sum((b[z, y, 1:] | b[z, y, :-1]).sum() for z in Z for y in Y)
          + sum((b[z, 1:, x] | b[z, :-1, x]).sum() for z in Z for x in X)
          + sum((b[1:, y, x] | b[:-1, y, x]).sum() for x in X for y in Y))
    return int(sv), int(se), int(sf), int(fv), int(fe), int(ev)


def np_betti(v):
    fg = v != 0
    b0 = ndimage.label(fg, structure=np.ones((3, 3, 3)))[1]
    lab, n = ndimage.label(~fg)
    border = set()
    for face in (lab[0], lab[-1], lab[:, 0], lab[:, -1], lab[:, :, 0], lab[:, :, -1]):
        border |= set(np.unique(face).tolist())
    border.discard(0)
    return b0, n - len(border)


def volumes():
    out = []
    rng = np.random.default_rng(11)
    for (d, h, w), frac, smooth in [((7, 9, 11), 0.3, 0), ((12, 10, 14), 0.5, 1),
                                    ((2, 2, 2), 0.5, 0), ((2, 3, 17), 0.6, 0),
                                    ((16, 16, 16), 0.45, 1), ((20, 5, 9), 0.7, 1)]:
        f = rng.standard_normal((d, h, w))
        if smooth:
            f = ndimage.uniform_filter(f, 3)
        out.append((f > np.quantile(f, 1 - frac)).astype(np.uint8))
    out.append(np.ones((4, 5, 6), np.uint8))
    out.append(np.zeros((4, 5, 6), np.uint8))
    return out


@pytest.mark.parametrize("idx", range(8))
def test_cubical_counts_match_whole_array_counts(idx):
    v = volumes()[idx]
    d, h, w = v.shape
    c = orc.cubical_measures(v, w, h, d)
    assert (c.vertices, c.edges, c.faces, c.cubes) == np_cubical(v)
    assert c.raw_euler == c.vertices - c.edges + c.faces - c.cubes


@pytest.mark.parametrize("idx", range(8))
def test_bonej_counts_match_whole_array_counts(idx):
    v = volumes()[idx]
    d, h, w = v.shape
    raw = orc.cubical_measures(v, w, h, d).raw_euler
    b = orc.bonej_connectivity(v, w, h, d, raw)
    sv, se, sf, fv, fe, ev = np_bonej_counts(v)
    f = sv
    e = se + 3 * f
    c = sf + 2 * e - 3 * f
    assert (b.f, b.e, b.c, b.d, b.a, b.b) == (f, e, c, ev + f, fv, fe)
    corr = (fv - fe + c) / 2.0 + (ev + f - e) / 4.0 + f / 8.0
    assert b.edge_correction == corr and b.connectivity == 1.0 - (raw - corr)


@pytest.mark.parametrize("idx", range(7))
def test_betti_numbers_match_scipy_labels(idx):
    v = volumes()[idx]
    d, h, w = v.shape
    raw = orc.cubical_measures(v, w, h, d).raw_euler
    t = orc.analyze_topology(v, w, h, d, raw)
    b0, b2 = np_betti(v)
    assert (t.beta0, t.beta2) == (b0, b2)
    assert t.beta1 == b0 + b2 - raw


def test_interface_and_roi_boundary_match_whole_array_counts():
    rng = np.random.default_rng(4)
    d, h, w = 9, 12, 15
    bone = (rng.random((d, h, w)) < 0.4).astype(np.uint8)
    roi = (rng.random((d, h, w)) < 0.8).astype(np.uint8)
    b, r = bone != 0, roi != 0
    want = 0
    for ax in range(3):
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[ax], sl1[ax] = slice(0, -1), slice(1, None)
        want += int((r[tuple(sl0)] & r[tuple(sl1)] & (b[tuple(sl0)] != b[tuple(sl1)])).sum())
    assert orc.interface_faces(bone, roi, w, h, d) == want
    ab = b & r
    pr = np.zeros((d + 2, h + 2, w + 2), bool)
    pr[1:-1, 1:-1, 1:-1] = r
    out = np.zeros_like(ab)
    for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        out |= ~pr[1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
    assert orc.roi_boundary_counts(ab.astype(np.uint8), roi, w, h, d) == (int(ab.sum()),
                                                                           int((ab & out).sum()))
