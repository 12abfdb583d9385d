"""Seeded synthetic trabecular volumes and ROI masks (libbm_synth.so).

volume(): white noise N(0,1) smoothed by a separable Gaussian (truncate 4
sigma, reflect boundary), thresholded so that round(bv_tv*N) voxels are bone;
bone = 255, marrow = 0, z-major layout.  Deterministic in its arguments.
"""
import numpy as np

from . import _native


def volume(width, height, depth, sigma, bv_tv, seed, threads=0):
    sx, sy, sz = (sigma, sigma, sigma) if np.isscalar(sigma) else sigma
    out = np.empty(width * height * depth, np.uint8)
    _native.synth().bm_synth_volume(out.ctypes.data, width, height, depth, float(sx), float(sy),
                                    float(sz), float(bv_tv), int(seed), int(threads))
    return out


def roi_full(width, height, depth):
    return np.ones(width * height * depth, np.uint8)


def roi_cylinder(width, height, depth, diameter, cx=None, cy=None):
    cx = (width - 1) / 2.0 if cx is None else cx
    cy = (height - 1) / 2.0 if cy is None else cy
    out = np.empty(width * height * depth, np.uint8)
    _native.synth().bm_synth_roi_cylinder(out.ctypes.data, width, height, depth, float(cx),
                                          float(cy), float(diameter) / 2.0)
    return out


def roi_ellipsoid(width, height, depth, semi_axes, center=None):
    c = ((width - 1) / 2.0, (height - 1) / 2.0, (depth - 1) / 2.0) if center is None else center
    out = np.empty(width * height * depth, np.uint8)
    _native.synth().bm_synth_roi_ellipsoid(out.ctypes.data, width, height, depth, float(c[0]),
                                           float(c[1]), float(c[2]), float(semi_axes[0]),
                                           float(semi_axes[1]), float(semi_axes[2]))
    return out
