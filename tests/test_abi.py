"""The C ABI library loads, exports every symbol include/bm.h declares, and
rejects invalid input with the reference's messages before touching a GPU."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import bone_morphometry_amd as bm
from bone_morphometry_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(bm_\w+)\(", text, re.M)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_version_and_empty_error():
    assert _native.lib().bm_version() == 100
    assert bm.require_isotropic_spacing((2.0, 2.0, 2.0)) == 2.0
    assert _native.lib().bm_last_error() == b""  # a call that succeeds clears it


# BinaryVolume::new, src/volume.rs:12-41
def test_binary_volume_new_rejects_bad_input():
    with pytest.raises(bm.BmError, match=r"volume has 7 voxels, but dimensions 2x2x2 require 8"):
        bm.BinaryVolume.new(np.zeros(7, np.uint8), 2, 2, 2)
    with pytest.raises(bm.BmError, match=r"at least 2 for 3-D morphometry; got 1x4x4"):
        bm.BinaryVolume.new(np.zeros(16, np.uint8), 1, 4, 4)
    bm.BinaryVolume.new(np.zeros(8, np.uint8), 2, 2, 2)


def test_calls_validate_before_device_work():
    v = bm.BinaryVolume(np.zeros(4, np.uint8), 1, 2, 2)
    with pytest.raises(bm.BmError) as e:
        bm.euclidean_distance_transform_squared(v, bm.Phase.Foreground)
    assert e.value.code == _native.BM_ERR_INVALID
    with pytest.raises(bm.BmError, match="at least 2"):
        bm.local_thickness(v, bm.Phase.Foreground)


def test_analyze_validates_spacing_and_shape():
    a = bm.BinaryVolume(np.zeros(8, np.uint8), 2, 2, 2)
    b = bm.BinaryVolume(np.zeros(12, np.uint8), 2, 2, 3)
    with pytest.raises(bm.BmError, match=r"voxel spacings must be finite and positive; got "
                                          r"\[1\.0, NaN, 2\.5\]"):
        bm.analyze_thickness_spacing(a, a, (1.0, math.nan, 2.5))
    with pytest.raises(bm.BmError, match="binary image dimensions 2x2x2 do not match ROI "
                                          "dimensions 2x2x3"):
        bm.analyze_thickness_spacing(a, b, (1.0, 1.0, 1.0))


def test_unsupported_sizes_are_reported():
    # checked before any allocation: 20000 > the per-axis limit
    big = _native.BmVolume(None, 0, 20000, 2, 2)
    rc = _native.lib().bm_euclidean_distance_transform_squared(ctypes.byref(big), 0, None, 0)
    assert rc == _native.BM_ERR_INVALID  # len 0 != 80000 fails BinaryVolume::new first
    rc = _native.lib().bm_euclidean_distance_transform_squared_device(None, 20000, 2, 2, 0, None,
                                                                      None)
    assert rc == _native.BM_ERR_UNSUPPORTED
    assert b"exceeds the device path's limits" in _native.lib().bm_last_error()


def test_tuning_cache_is_parsed_without_a_device(tmp_path):
    good = tmp_path / "good.json"
    good.write_text('{"gfx950": {"h384_d320": {"y_rows": 256, "y_halo": 32, "z_rows": 512, '
                    '"z_halo": 64, "fill_cap": 512, "edt_y_ms": 0.1}}}')
    bm.tuning_load(good)
    bad = tmp_path / "bad.json"
    bad.write_text('{"gfx950": {"h384_d320": {"y_rows": }}}')
    with pytest.raises(bm.BmError) as e:
        bm.tuning_load(bad)
    assert e.value.code == _native.BM_ERR_INVALID and "tuning cache" in str(e.value)
    bm.tuning_load(tmp_path / "absent.json")  # no file: the defaults


# require_isotropic_spacing, src/sdf_curvature.rs:147-160 (host side, no device)
def test_require_isotropic_spacing_messages():
    assert bm.require_isotropic_spacing((10.5, 10.5, 10.5)) == 10.5
    assert bm.require_isotropic_spacing((8.0, 8.0 + 5e-9, 8.0 - 5e-9)) == 8.0  # within 1e-9*h
    with pytest.raises(bm.BmError, match="^voxel spacing must be finite and positive$"):
        bm.require_isotropic_spacing((1.0, math.nan, 1.0))
    with pytest.raises(bm.BmError, match="voxel spacing must be finite and positive"):
        bm.require_isotropic_spacing((0.0, 1.0, 1.0))
    # Rust's Display of f64: no exponent, shortest digits, no trailing ".0"
    with pytest.raises(bm.BmError) as e:
        bm.require_isotropic_spacing((1.0, 1.5, 1.0))
    assert str(e.value) == ("advanced SDF curvature currently requires isotropic spacing; "
                            "got [1, 1.5, 1] µm")
    with pytest.raises(bm.BmError) as e:
        bm.require_isotropic_spacing((1e-7, 2.5e-7, 1e21))
    assert str(e.value).endswith("got [0.0000001, 0.00000025, 1000000000000000000000] µm")


# gaussian_kernel, src/sdf_curvature.rs:182-198 (host side): the library's
# weights are the oracle's to the bit
def test_gaussian_kernel_matches_oracle():
    import _oracle as orc
    for sigma in (0.125, 0.3, 0.5, 1.0, 1.5, 2.0, 2.857142857142857, 3.75, 7.1, 31.0, 120.5):
        got = bm.gaussian_kernel(sigma)
        want = orc.gaussian_kernel(sigma)
        assert got.size == want.size and np.array_equal(got.view(np.uint64),
                                                         want.view(np.uint64)), sigma
    assert bm.gaussian_kernel(0.0).tolist() == [1.0]
    assert math.isnan(bm.gaussian_kernel(math.nan)[0])
    with pytest.raises(bm.BmError) as e:
        bm.gaussian_kernel(1e7)
    assert e.value.code == _native.BM_ERR_UNSUPPORTED
