"""The analyze entry point that allocates nothing
(bm_analyze_thickness_spacing_device_noalloc): through the C ABI, its results
equal bm_analyze_thickness_spacing_device's bit for bit on two calls in a row,
the second on a smaller volume (the device scratch keeps the first call's
contents), and sizes beyond its scratch are rejected with BM_ERR_UNSUPPORTED
before any device work."""
import ctypes

import numpy as np
import pytest

import bone_morphometry_amd as bm
from bone_morphometry_amd import _native, hip, synth

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 256 << 20


def _lib():
    L = _native.lib()
    f = L.bm_analyze_thickness_spacing_device_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                  ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                  ctypes.POINTER(_native.BmThicknessSpacing), ctypes.c_void_p]
    f.restype = ctypes.c_int
    return L


def _noalloc(b, r, w, h, d, spacing, mode=bm.FeatureMode.Standard):
    L = _lib()
    out = _native.BmThicknessSpacing()
    sp = (ctypes.c_double * 3)(*spacing)
    rc = L.bm_analyze_thickness_spacing_device_noalloc(b, r, w, h, d, sp, int(mode),
                                                        ctypes.byref(out), None)
    if rc != _native.BM_OK:
        raise bm.BmError(rc, L.bm_last_error().decode())
    return bm.ThicknessSpacing._from_c(out)


def _fields(ts):
    out = []
    for s in (ts.thickness, ts.spacing):
        out += [s.mean, s.standard_deviation, s.maximum, s.median, s.percentile_10,
                s.percentile_25, s.percentile_75, s.percentile_90, s.interquartile_range,
                s.coefficient_of_variation, s.count]
    return out + [ts.bone_volume, ts.total_volume, ts.bone_volume_fraction]


def _volume(w, h, d, sigma, bvtv, seed, frac, sp):
    bone = synth.volume(w, h, d, sigma, bvtv, seed)
    roi = synth.roi_cylinder(w, h, d, frac * min(w, h))
    b, r = hip.DeviceBuffer(bone.size), hip.DeviceBuffer(roi.size)
    b.upload(bone)
    r.upload(roi)
    return b, r, (sp, sp, sp)


def test_noalloc_matches_allocating_call_twice():
    shapes = [((190, 176, 200), 4.0, 0.2, 41, 0.9, 11.0),  # near the scratch limit
              ((97, 64, 81), 2.5, 0.3, 42, 0.8, 7.5)]      # then smaller
    for (w, h, d), sigma, bvtv, seed, frac, sp in shapes:
        b, r, spacing = _volume(w, h, d, sigma, bvtv, seed, frac, sp)
        want = bm.api.analyze_thickness_spacing_device(b.value, r.value, w, h, d, spacing)
        got = _noalloc(b.value, r.value, w, h, d, spacing)
        np.testing.assert_array_equal(np.array(_fields(got)), np.array(_fields(want)))
        # the Classic mode (no quantiles) through the same scratch
        want = bm.api.analyze_thickness_spacing_device(b.value, r.value, w, h, d, spacing,
                                                       bm.FeatureMode.Classic)
        got = _noalloc(b.value, r.value, w, h, d, spacing, bm.FeatureMode.Classic)
        np.testing.assert_array_equal(np.array(_fields(got)), np.array(_fields(want)))


def test_noalloc_rejects_volumes_beyond_its_scratch():
    b, r, spacing = _volume(16, 16, 16, 1.5, 0.3, 43, 0.9, 10.0)
    with pytest.raises(bm.BmError, match="bytes of device scratch") as e:
        _noalloc(b.value, r.value, 512, 512, 512, spacing)
    assert e.value.code == _native.BM_ERR_UNSUPPORTED
    with pytest.raises(bm.BmError, match="ROI contains no bone voxels"):
        z = hip.DeviceBuffer(16 ** 3)
        z.upload(np.zeros(16 ** 3, np.uint8))
        _noalloc(z.value, r.value, 16, 16, 16, spacing)
