"""The device form of the input step: TIFF pages decoded on host threads and
streamed through pinned staging buffers into HBM equal the host read, byte for
byte, also when the volume spans several staging batches."""
import numpy as np
import pytest
from PIL import Image

import bone_morphometry_amd as bm
from bone_morphometry_amd import hip, synth

pytestmark = pytest.mark.gpu


def save(path, vol, compression):
    pages = [Image.fromarray(p) for p in vol]
    pages[0].save(path, save_all=True, append_images=pages[1:], compression=compression)


@pytest.mark.parametrize("shape,compression", [((9, 40, 70), "tiff_lzw"),
                                               ((300, 512, 512), "raw"),
                                               ((140, 384, 400), "tiff_adobe_deflate")])
def test_device_read_equals_host_read(tmp_path, shape, compression):
    d, h, w = shape
    vol = synth.volume(w, h, d, 2.0, 0.3, 5).reshape(shape)
    path = str(tmp_path / "v.tif")
    save(path, vol, compression)
    host = bm.read_binary_tiff(path, strict_binary=True)
    buf = hip.DeviceBuffer(w * h * d)
    dims = bm.read_binary_tiff_device(path, buf.value, w * h * d, strict_binary=True)
    assert dims == (w, h, d)
    got = np.empty(w * h * d, np.uint8)
    buf.download(got)
    assert np.array_equal(got, host.data)
    assert np.array_equal(got, (vol > 0).astype(np.uint8).reshape(-1))


def test_device_roi_read_and_errors(tmp_path):
    roi = synth.roi_cylinder(64, 48, 20, 40.0).reshape(20, 48, 64)
    path = str(tmp_path / "roi.tif")
    save(path, roi, "raw")
    buf = hip.DeviceBuffer(roi.size)
    assert bm.read_roi_tiff_device(path, buf.value, roi.size) == (64, 48, 20)
    got = np.empty(roi.size, np.uint8)
    buf.download(got)
    assert np.array_equal(got, (roi > 0).astype(np.uint8).reshape(-1))
    roi[7, 1, 2] = 3
    save(path, roi, "raw")
    with pytest.raises(bm.BmError, match=r"page 8, sample 66 has value 3 \(expected 0/1 or 0/255\)"):
        bm.read_roi_tiff_device(path, buf.value, roi.size)
    with pytest.raises(bm.BmError, match="output buffer holds 10 bytes"):
        bm.read_roi_tiff_device(path, buf.value, 10)


def test_tiff_to_analyze(tmp_path):
    # the input step feeding the path: files -> HBM -> analyze, equal to the host route
    w, h, d = 96, 80, 60
    bone = synth.volume(w, h, d, 2.5, 0.2, 11).reshape(d, h, w)
    roi = synth.roi_cylinder(w, h, d, 70.0).reshape(d, h, w)
    bp, rp = str(tmp_path / "bone.tif"), str(tmp_path / "roi.tif")
    save(bp, bone, "tiff_lzw")
    save(rp, roi, "packbits")
    db, dr = hip.DeviceBuffer(bone.size), hip.DeviceBuffer(roi.size)
    bm.read_binary_tiff_device(bp, db.value, bone.size, strict_binary=True)
    bm.read_roi_tiff_device(rp, dr.value, roi.size)
    m = bm.analyze_device(db.value, dr.value, w, h, d, (12.0, 12.0, 12.0))
    ref = bm.analyze(bm.read_binary_tiff(bp, strict_binary=True), bm.read_roi_tiff(rp),
                     (12.0, 12.0, 12.0))
    assert m == ref
