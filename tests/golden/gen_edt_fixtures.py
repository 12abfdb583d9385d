"""Writes tests/golden/edt_scipy.npz: squared-EDT fixtures from scipy.

For each case the selected phase is `vol != 0`; scipy's
distance_transform_edt(selected, return_indices=True) gives, for every
selected voxel, the index of a nearest non-selected voxel inside the array
(no exterior halo, the convention of src/edt.rs:51-56), and the squared
integer offset to it is the exact squared EDT.  Run here (scipy is available
in the build container); the GPU host only reads the .npz.

Shapes cover odd sizes, rows longer than one 512-voxel segment (X pass) and
columns/lines longer than one LDS tile (Y and Z passes); densities range from
a few isolated sites to dense sites.
"""
import os

import numpy as np
import scipy.ndimage as ndi

HERE = os.path.dirname(os.path.abspath(__file__))

# (width, height, depth, site probability)
CASES = [
    (2, 2, 2, 0.5), (3, 4, 5, 0.3), (7, 5, 3, 0.1), (16, 16, 16, 0.05), (17, 9, 33, 0.2),
    (33, 31, 29, 0.01), (64, 64, 8, 0.002), (1100, 3, 2, 0.004), (600, 5, 4, 0.01),
    (5, 700, 3, 0.003), (4, 3, 1300, 0.002), (20, 12, 650, 0.0005), (13, 530, 11, 0.0008),
]


def squared_edt(sel):
    if not (~sel).any():
        raise ValueError("all selected")
    _, idx = ndi.distance_transform_edt(sel, return_indices=True)
    grid = np.indices(sel.shape)
    d2 = ((idx - grid).astype(np.int64) ** 2).sum(axis=0)
    d2[~sel] = 0
    return d2.astype(np.uint32)


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for k, (w, h, d, p) in enumerate(CASES):
        sites = rng.random((d, h, w)) < p
        if not sites.any():
            sites.flat[rng.integers(sites.size)] = True
        vol = (~sites).astype(np.uint8) * 255
        out[f"vol_{k}"] = vol.reshape(-1)
        out[f"shape_{k}"] = np.array([w, h, d], np.int64)
        out[f"d2_{k}"] = squared_edt(vol != 0).reshape(-1)
    np.savez_compressed(os.path.join(HERE, "edt_scipy.npz"), **out)


if __name__ == "__main__":
    main()
