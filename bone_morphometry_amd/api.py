"""Reference-shaped calls over libbm.so (see the package docstring)."""
import ctypes
import os
import dataclasses
import enum

import numpy as np

from . import _native


class BmError(RuntimeError):
    """A failed call; .code is the BM_ERR_* value, str() the reference's message."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def _check(rc):
    if rc != _native.BM_OK:
        raise BmError(rc, _native.lib().bm_last_error().decode())


class Phase(enum.IntEnum):
    """src/edt.rs:8-24."""
    Foreground = 0
    Background = 1

    def is_selected(self, value):
        return value != 0 if self is Phase.Foreground else value == 0


class FeatureMode(enum.IntEnum):
    """src/metrics.rs:12-24."""
    Bonej = 0
    Standard = 1
    Classic = 2
    Refined = 3
    Experimental = 4


class BinaryVolume:
    """src/volume.rs:3-9: u8 voxels, index (z*height + y)*width + x."""

    def __init__(self, data, width, height, depth):
        self.data = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
        self.width, self.height, self.depth = int(width), int(height), int(depth)

    @classmethod
    def new(cls, data, width, height, depth):
        """BinaryVolume::new, src/volume.rs:12-41 (validation by the library)."""
        v = cls(data, width, height, depth)
        vol = v._c()
        _check(_native.lib().bm_binary_volume_check(ctypes.byref(vol)))
        return v

    def __len__(self):
        return self.data.size

    def index(self, x, y, z):
        return (z * self.height + y) * self.width + x

    def _c(self):
        return _native.BmVolume(self.data.ctypes.data, self.data.size, self.width, self.height,
                                self.depth)

    def phase_inside(self, roi, select_bone):
        """src/volume.rs:84-105, computed on the device."""
        out = np.empty(self.data.size, np.uint8)
        a, b = self._c(), roi._c()
        _check(_native.lib().bm_phase_inside(ctypes.byref(a), ctypes.byref(b), int(select_bone),
                                             out.ctypes.data, out.size))
        return BinaryVolume(out, self.width, self.height, self.depth)


def euclidean_distance_transform_squared(volume, phase):
    """src/edt.rs:57-156 -> np.ndarray[uint32] of len(volume)."""
    out = np.empty(volume.data.size, np.uint32)
    v = volume._c()
    _check(_native.lib().bm_euclidean_distance_transform_squared(
        ctypes.byref(v), int(phase), out.ctypes.data, out.size))
    return out


def local_thickness(volume, phase):
    """src/local_thickness.rs:16-44 -> np.ndarray[float32] of diameters."""
    out = np.empty(volume.data.size, np.float32)
    v = volume._c()
    _check(_native.lib().bm_local_thickness(ctypes.byref(v), int(phase), out.ctypes.data,
                                            out.size))
    return out


def local_thickness_stages(volume, phase):
    """Squared EDT, ridge flags, propagated r^2 map and final map of local_thickness."""
    n = volume.data.size
    sq = np.empty(n, np.uint32)
    flags = np.empty(n, np.uint8)
    prop = np.empty(n, np.uint32)
    th = np.empty(n, np.float32)
    rc, fv = ctypes.c_uint64(0), ctypes.c_uint64(0)
    v = volume._c()
    _check(_native.lib().bm_local_thickness_stages(
        ctypes.byref(v), int(phase), sq.ctypes.data, flags.ctypes.data, prop.ctypes.data,
        th.ctypes.data, ctypes.byref(rc), ctypes.byref(fv)))
    return dict(squared_edt=sq, ridge_flags=flags, propagated=prop, thickness=th,
                ridge_count=rc.value, fill_visits=fv.value)


class StatsOrder(enum.IntEnum):
    """Order of the f64 sums behind mean and SD (BM_STATS_* in bm.h)."""
    Exact = 0       # exact device sums, rounded once
    Sequential = 1  # the reference's order (src/metrics.rs:215-226), bit for bit


def set_stats_order(order):
    """Sets the process-wide statistics order; returns the previous one."""
    prev = ctypes.c_int(0)
    _check(_native.lib().bm_set_stats_order(int(order), ctypes.byref(prev)))
    return StatsOrder(prev.value)


class stats_order:
    """Context manager: `with stats_order(StatsOrder.Sequential): ...`."""

    def __init__(self, order):
        self.order = order
        self.prev = None

    def __enter__(self):
        self.prev = set_stats_order(self.order)
        return self

    def __exit__(self, *exc):
        set_stats_order(self.prev)
        return False


def require_isotropic_spacing(spacing):
    """src/sdf_curvature.rs:147-160 -> h (the common spacing) or BmError."""
    sp = (ctypes.c_double * 3)(*[float(v) for v in spacing])
    h = ctypes.c_double(0.0)
    _check(_native.lib().bm_require_isotropic_spacing(sp, ctypes.byref(h)))
    return h.value


def gaussian_kernel(sigma):
    """src/sdf_curvature.rs:182-198 -> np.ndarray[float64] of 2r+1 normalised weights."""
    n = ctypes.c_uint64(0)
    _check(_native.lib().bm_gaussian_kernel(ctypes.c_double(sigma), None, 0, ctypes.byref(n)))
    out = np.empty(n.value, np.float64)
    _check(_native.lib().bm_gaussian_kernel(ctypes.c_double(sigma), out.ctypes.data, out.size,
                                            ctypes.byref(n)))
    return out


def signed_distance_field_um(bone, h):
    """src/sdf_curvature.rs:162-180 -> np.ndarray[float64]: +sqrt(d2) h inside, -sqrt(d2') h out."""
    out = np.empty(bone.data.size, np.float64)
    v = bone._c()
    _check(_native.lib().bm_signed_distance_field_um(ctypes.byref(v), ctypes.c_double(h),
                                                     out.ctypes.data, out.size))
    return out


def gaussian_smooth_nearest_in_place(data, width, height, depth, sigma_vox):
    """src/sdf_curvature.rs:200-269 on a C-contiguous float64 array, in place."""
    if not (isinstance(data, np.ndarray) and data.dtype == np.float64
            and data.flags.c_contiguous and data.flags.writeable):
        raise TypeError("data must be a writeable C-contiguous float64 array")
    _check(_native.lib().bm_gaussian_smooth_nearest_in_place(
        data.ctypes.data, data.size, int(width), int(height), int(depth),
        ctypes.c_double(sigma_vox)))
    return data


def smoothed_signed_distance_field_um(bone, h, sigma_vox):
    """signed_distance_field_um then gaussian_smooth_nearest_in_place, as the callers at
    src/sdf_curvature.rs:98-105 and src/minkowski.rs:64-71 make them."""
    out = np.empty(bone.data.size, np.float64)
    v = bone._c()
    _check(_native.lib().bm_smoothed_signed_distance_field_um(
        ctypes.byref(v), ctypes.c_double(h), ctypes.c_double(sigma_vox), out.ctypes.data,
        out.size))
    return out


def signed_distance_field_um_device(d_bone_ptr, width, height, depth, h, d_out_ptr, stream=None):
    """Device-resident form (device pointers, e.g. torch tensor.data_ptr())."""
    _check(_native.lib().bm_signed_distance_field_um_device(
        d_bone_ptr, width, height, depth, ctypes.c_double(h), d_out_ptr, stream))


def gaussian_smooth_nearest_in_place_device(d_data_ptr, width, height, depth, sigma_vox,
                                            stream=None):
    """Device-resident form: smooths width*height*depth float64 values in place."""
    _check(_native.lib().bm_gaussian_smooth_nearest_in_place_device(
        d_data_ptr, width, height, depth, ctypes.c_double(sigma_vox), stream))


def smoothed_signed_distance_field_um_device(d_bone_ptr, width, height, depth, h, sigma_vox,
                                             d_out_ptr, stream=None):
    """Device-resident form of smoothed_signed_distance_field_um."""
    _check(_native.lib().bm_smoothed_signed_distance_field_um_device(
        d_bone_ptr, width, height, depth, ctypes.c_double(h), ctypes.c_double(sigma_vox),
        d_out_ptr, stream))


@dataclasses.dataclass
class DistributionStats:
    """src/metrics.rs:26-38 (+ count)."""
    mean: float
    standard_deviation: float
    maximum: float
    median: float
    percentile_10: float
    percentile_25: float
    percentile_75: float
    percentile_90: float
    interquartile_range: float
    coefficient_of_variation: float
    count: int

    @classmethod
    def _from_c(cls, s):
        return cls(**{f.name: getattr(s, f.name) for f in dataclasses.fields(cls)})


@dataclasses.dataclass
class ThicknessSpacing:
    thickness: DistributionStats
    spacing: DistributionStats
    bone_volume: float
    total_volume: float
    bone_volume_fraction: float

    @classmethod
    def _from_c(cls, r):
        return cls(DistributionStats._from_c(r.thickness), DistributionStats._from_c(r.spacing),
                   r.bone_volume, r.total_volume, r.bone_volume_fraction)


def analyze_thickness_spacing(bone, roi, voxel_spacing, feature_mode=FeatureMode.Standard):
    """Thickness/spacing part of analyze, src/metrics.rs:64-132."""
    sp = (ctypes.c_double * 3)(*[float(v) for v in voxel_spacing])
    out = _native.BmThicknessSpacing()
    a, b = bone._c(), roi._c()
    _check(_native.lib().bm_analyze_thickness_spacing(ctypes.byref(a), ctypes.byref(b), sp,
                                                      int(feature_mode), ctypes.byref(out)))
    return ThicknessSpacing._from_c(out)


def analyze_thickness_spacing_device(d_bone_ptr, d_roi_ptr, width, height, depth, voxel_spacing,
                                     feature_mode=FeatureMode.Standard, stream=None):
    """Device-resident form: bone/roi are device pointers (e.g. torch tensor.data_ptr())."""
    sp = (ctypes.c_double * 3)(*[float(v) for v in voxel_spacing])
    out = _native.BmThicknessSpacing()
    _check(_native.lib().bm_analyze_thickness_spacing_device(
        d_bone_ptr, d_roi_ptr, width, height, depth, sp, int(feature_mode), ctypes.byref(out),
        stream))
    return ThicknessSpacing._from_c(out)


# ---------------------------------------------------------------------------
# The rest of analyze (src/topology.rs, src/metrics.rs:134-202).

def _struct(cls, c):
    return cls(**{f.name: getattr(c, f.name) for f in dataclasses.fields(cls)})


@dataclasses.dataclass
class CubicalMeasures:
    """src/topology.rs:13-17."""
    raw_euler: int
    mean_breadth_lattice_units: float


@dataclasses.dataclass
class Topology:
    """src/topology.rs:5-11."""
    raw_euler: int
    beta0: int
    beta1: int
    beta2: int


@dataclasses.dataclass
class BonejConnectivity:
    """src/topology.rs:19-27."""
    edge_correction: float
    delta_chi: float
    connectivity: float


@dataclasses.dataclass
class GeometryMetrics:
    """src/metrics.rs:40-46."""
    bone_surface: float
    bone_surface_to_bone_volume: float
    bone_surface_to_total_volume: float
    mean_breadth: float


@dataclasses.dataclass
class Metrics:
    """src/metrics.rs:48-62."""
    thickness: DistributionStats
    spacing: DistributionStats
    bone_volume: float
    total_volume: float
    bone_volume_fraction: float
    topology: Topology
    bonej_connectivity: BonejConnectivity
    bonej_connectivity_density: float
    connectivity: float
    connectivity_density: float
    roi_boundary_bone_fraction: float
    geometry: GeometryMetrics

    @classmethod
    def _from_c(cls, m):
        return cls(DistributionStats._from_c(m.thickness), DistributionStats._from_c(m.spacing),
                   m.bone_volume, m.total_volume, m.bone_volume_fraction,
                   _struct(Topology, m.topology), _struct(BonejConnectivity, m.bonej_connectivity),
                   m.bonej_connectivity_density, m.connectivity, m.connectivity_density,
                   m.roi_boundary_bone_fraction, _struct(GeometryMetrics, m.geometry))


def cubical_measures(volume):
    """src/topology.rs:53-171 on the volume's nonzero voxels."""
    out = _native.BmCubical()
    v = volume._c()
    _check(_native.lib().bm_cubical_measures(ctypes.byref(v), ctypes.byref(out)))
    return _struct(CubicalMeasures, out)


def analyze_topology(volume, raw_euler):
    """src/topology.rs:29-51: Betti numbers of the volume's nonzero voxels."""
    out = _native.BmTopology()
    v = volume._c()
    _check(_native.lib().bm_analyze_topology(ctypes.byref(v), int(raw_euler), ctypes.byref(out)))
    return _struct(Topology, out)


def bonej_connectivity(volume, raw_euler):
    """src/topology.rs:175-201: BoneJ's stack edge correction."""
    out = _native.BmBonej()
    v = volume._c()
    _check(_native.lib().bm_bonej_connectivity(ctypes.byref(v), int(raw_euler),
                                               ctypes.byref(out)))
    return _struct(BonejConnectivity, out)


def bone_marrow_interface_faces(bone, roi):
    """src/topology.rs:442-490."""
    out = ctypes.c_int64(0)
    b, r = bone._c(), roi._c()
    _check(_native.lib().bm_bone_marrow_interface_faces(ctypes.byref(b), ctypes.byref(r),
                                                        ctypes.byref(out)))
    return out.value


def analyze(bone, roi, voxel_spacing, feature_mode=FeatureMode.Standard):
    """analyze, src/metrics.rs:64-164: every field of Metrics."""
    sp = (ctypes.c_double * 3)(*[float(v) for v in voxel_spacing])
    out = _native.BmMetrics()
    b, r = bone._c(), roi._c()
    _check(_native.lib().bm_analyze(ctypes.byref(b), ctypes.byref(r), sp, int(feature_mode),
                                    ctypes.byref(out)))
    return Metrics._from_c(out)


def analyze_device(d_bone_ptr, d_roi_ptr, width, height, depth, voxel_spacing,
                   feature_mode=FeatureMode.Standard, stream=None):
    """Device-resident analyze (device pointers, e.g. from hip.DeviceBuffer)."""
    sp = (ctypes.c_double * 3)(*[float(v) for v in voxel_spacing])
    out = _native.BmMetrics()
    _check(_native.lib().bm_analyze_device(d_bone_ptr, d_roi_ptr, width, height, depth, sp,
                                           int(feature_mode), ctypes.byref(out), stream))
    return Metrics._from_c(out)


# ---------------------------------------------------------------------------
# The input step (src/tiff_io.rs).

def _path(path):
    return os.fsencode(path)


def _read_tiff(call, path, *args):
    dims = (ctypes.c_uint64 * 3)()
    _check(call(_path(path), *args, None, 0, dims))
    w, h, d = dims
    out = np.empty(w * h * d, np.uint8)
    _check(call(_path(path), *args, out.ctypes.data, out.size, dims))
    return BinaryVolume(out, w, h, d)


def read_binary_tiff(path, threshold=0, strict_binary=False, threads=0):
    """src/tiff_io.rs:18-26: a multi-page 8-bit TIFF as 0/1 voxels (value > threshold)."""
    return _read_tiff(_native.lib().bm_read_binary_tiff, path, int(threshold), int(strict_binary),
                      int(threads))


def read_roi_tiff(path, threads=0):
    """src/tiff_io.rs:28-31: a 0/1 or 0/255 ROI mask."""
    return _read_tiff(_native.lib().bm_read_roi_tiff, path, int(threads))


def read_binary_tiff_device(path, d_out_ptr, capacity, threshold=0, strict_binary=False,
                            threads=0, stream=None):
    """Device form: the volume decoded into d_out_ptr; returns (width, height, depth)."""
    dims = (ctypes.c_uint64 * 3)()
    _check(_native.lib().bm_read_binary_tiff_device(_path(path), int(threshold),
                                                    int(strict_binary), int(threads), d_out_ptr,
                                                    capacity, dims, stream))
    return tuple(dims)


def read_roi_tiff_device(path, d_out_ptr, capacity, threads=0, stream=None):
    """Device form of read_roi_tiff; returns (width, height, depth)."""
    dims = (ctypes.c_uint64 * 3)()
    _check(_native.lib().bm_read_roi_tiff_device(_path(path), int(threads), d_out_ptr, capacity,
                                                 dims, stream))
    return tuple(dims)


def read_isotropic_spacing_um(path):
    """src/tiff_io.rs:149-317: [x, y, z] voxel spacing in micrometres."""
    out = (ctypes.c_double * 3)()
    _check(_native.lib().bm_read_isotropic_spacing_um(_path(path), out))
    return [out[0], out[1], out[2]]


# ---------------------------------------------------------------------------
# Batched and serving forms (include/bm.h: bm_batch_*, bm_server_*).

def _batch_items(items):
    """items: (d_bone_ptr, d_roi_ptr, width, height, depth, voxel_spacing, feature_mode)."""
    arr = (_native.BmBatchItem * len(items))()
    for i, (b, r, w, h, d, sp, mode) in enumerate(items):
        arr[i].d_bone, arr[i].d_roi = b, r
        arr[i].width, arr[i].height, arr[i].depth = int(w), int(h), int(d)
        for k in range(3):
            arr[i].voxel_spacing[k] = float(sp[k])
        arr[i].feature_mode = int(mode)
    return arr


def _result(r):
    if r.status != _native.BM_OK:
        return BmError(r.status, r.message.decode())
    return ThicknessSpacing._from_c(r.metrics)


class Batch:
    """Many device-resident analyze() calls as one plan (one launch per stage),
    optionally captured once into a hipGraph.  run() returns, per volume, a
    ThicknessSpacing or the BmError the one-volume call would raise."""

    def __init__(self, items, use_graph=True):
        self._items = _batch_items(items)
        self.count = len(items)
        self._h = ctypes.c_void_p()
        _check(_native.lib().bm_batch_create(self._items, self.count, 1 if use_graph else 0,
                                             ctypes.byref(self._h)))
        self._res = (_native.BmBatchResult * self.count)()

    def run(self):
        _check(_native.lib().bm_batch_run(self._h, self._res))
        return [_result(self._res[i]) for i in range(self.count)]

    def device_bytes(self):
        v = ctypes.c_uint64()
        _check(_native.lib().bm_batch_device_bytes(self._h, ctypes.byref(v)))
        return v.value

    def close(self):
        if self._h:
            _check(_native.lib().bm_batch_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(items, cache_path, verbose=False):
    """Times the tunables on the largest volume of each shape class of items
    (device-resident, as for Batch) and writes them to the JSON cache."""
    arr = _batch_items(items)
    _check(_native.lib().bm_autotune(arr, len(items), str(cache_path).encode(),
                                     1 if verbose else 0))


def tuning_load(cache_path):
    _check(_native.lib().bm_tuning_load(str(cache_path).encode()))


class Server:
    """The request loop of the serving form: submit() host-resident volumes,
    poll() their results with per-request timestamps (seconds since the
    server started).  The arrays passed to submit() are kept referenced until
    their request is polled done."""

    def __init__(self, budget_bytes, max_batch=32, slots=2):
        self._h = ctypes.c_void_p()
        _check(_native.lib().bm_server_create(int(budget_bytes), int(max_batch), int(slots),
                                              ctypes.byref(self._h)))
        self._keep = {}

    def submit(self, bone, roi, width, height, depth, voxel_spacing,
               feature_mode=FeatureMode.Standard):
        bone = np.ascontiguousarray(bone, dtype=np.uint8).reshape(-1)
        roi = np.ascontiguousarray(roi, dtype=np.uint8).reshape(-1)
        sp = (ctypes.c_double * 3)(*[float(v) for v in voxel_spacing])
        t = ctypes.c_uint64()
        _check(_native.lib().bm_server_submit(self._h, bone.ctypes.data, roi.ctypes.data,
                                              int(width), int(height), int(depth), sp,
                                              int(feature_mode), ctypes.byref(t)))
        self._keep[t.value] = (bone, roi)
        return t.value

    def poll(self, ticket, wait=True):
        """None while running, else (ThicknessSpacing or BmError, timing dict)."""
        res = _native.BmBatchResult()
        tim = _native.BmRequestTiming()
        done = ctypes.c_int(0)
        _check(_native.lib().bm_server_poll(self._h, int(ticket), 1 if wait else 0,
                                            ctypes.byref(res), ctypes.byref(tim),
                                            ctypes.byref(done)))
        if not done.value:
            return None
        self._keep.pop(ticket, None)
        return _result(res), dict(submit_s=tim.submit_s, start_s=tim.start_s, done_s=tim.done_s,
                                  batch_size=tim.batch_size)

    def cancel(self, ticket):
        """Withdraws a still-queued request (True); False once it has launched."""
        c = ctypes.c_int(0)
        _check(_native.lib().bm_server_cancel(self._h, int(ticket), ctypes.byref(c)))
        return bool(c.value)

    def stats(self):
        s = _native.BmServerStats()
        _check(_native.lib().bm_server_stats_get(self._h, ctypes.byref(s)))
        return {f[0]: getattr(s, f[0]) for f in s._fields_}

    def close(self):
        if self._h:
            _check(_native.lib().bm_server_destroy(self._h))
            self._h = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
