"""ctypes binding of include/bm.h (libbm.so) and of the synthetic generator.

The library is loaded from this package directory.  It is the only compute path:
there is no CPU fallback, and a missing or unloadable library raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# BM_LIBRARY names another build of the same library (the host-sanitized
# build of the CPU tests); by default the one built in this directory.
LIB_PATH = os.environ.get("BM_LIBRARY") or os.path.join(HERE, "libbm.so")
SYNTH_PATH = os.path.join(HERE, "libbm_synth.so")
# The autotuner's cache shipped with the package (per architecture and shape
# class); BM_TUNE_CACHE, when set, takes precedence.
TUNE_PATH = os.path.join(HERE, "tuning.json")

BM_OK = 0
BM_ERR_INVALID = 1
BM_ERR_DEVICE = 2
BM_ERR_UNSUPPORTED = 3

# Every symbol include/bm.h declares.
EXPORTS = (
    "bm_last_error", "bm_version", "bm_device_count", "bm_binary_volume_check", "bm_phase_inside",
    "bm_euclidean_distance_transform_squared", "bm_local_thickness",
    "bm_analyze_thickness_spacing", "bm_euclidean_distance_transform_squared_device",
    "bm_local_thickness_device", "bm_analyze_thickness_spacing_device",
    "bm_analyze_thickness_spacing_device_noalloc",
    "bm_local_thickness_stages", "bm_set_stats_order", "bm_profile_enable", "bm_profile_read",
    "bm_profile_only",
    "bm_batch_create", "bm_batch_run", "bm_batch_device_bytes", "bm_batch_destroy",
    "bm_autotune", "bm_tuning_load",
    "bm_server_create", "bm_server_submit", "bm_server_poll", "bm_server_cancel",
    "bm_server_stats_get",
    "bm_server_destroy",
    "bm_require_isotropic_spacing", "bm_gaussian_kernel", "bm_signed_distance_field_um",
    "bm_gaussian_smooth_nearest_in_place", "bm_smoothed_signed_distance_field_um",
    "bm_signed_distance_field_um_device", "bm_gaussian_smooth_nearest_in_place_device",
    "bm_smoothed_signed_distance_field_um_device",
    "bm_cubical_measures", "bm_analyze_topology", "bm_bonej_connectivity",
    "bm_bone_marrow_interface_faces", "bm_analyze", "bm_analyze_device",
    "bm_read_binary_tiff", "bm_read_roi_tiff", "bm_read_binary_tiff_device",
    "bm_read_roi_tiff_device", "bm_read_isotropic_spacing_um",
)


class BmVolume(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("len", ctypes.c_uint64), ("width", ctypes.c_uint64),
                ("height", ctypes.c_uint64), ("depth", ctypes.c_uint64)]


class BmDistributionStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "mean", "standard_deviation", "maximum", "median", "percentile_10", "percentile_25",
        "percentile_75", "percentile_90", "interquartile_range", "coefficient_of_variation")] + [
        ("count", ctypes.c_uint64)]


class BmThicknessSpacing(ctypes.Structure):
    _fields_ = [("thickness", BmDistributionStats), ("spacing", BmDistributionStats),
                ("bone_volume", ctypes.c_double), ("total_volume", ctypes.c_double),
                ("bone_volume_fraction", ctypes.c_double)]


class BmCubical(ctypes.Structure):
    _fields_ = [("raw_euler", ctypes.c_int64), ("mean_breadth_lattice_units", ctypes.c_double)]


class BmTopology(ctypes.Structure):
    _fields_ = [("raw_euler", ctypes.c_int64), ("beta0", ctypes.c_uint64),
                ("beta1", ctypes.c_uint64), ("beta2", ctypes.c_uint64)]


class BmBonej(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("edge_correction", "delta_chi", "connectivity")]


class BmGeometry(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "bone_surface", "bone_surface_to_bone_volume", "bone_surface_to_total_volume",
        "mean_breadth")]


class BmMetrics(ctypes.Structure):
    _fields_ = [("thickness", BmDistributionStats), ("spacing", BmDistributionStats),
                ("bone_volume", ctypes.c_double), ("total_volume", ctypes.c_double),
                ("bone_volume_fraction", ctypes.c_double), ("topology", BmTopology),
                ("bonej_connectivity", BmBonej)] + [(n, ctypes.c_double) for n in (
                    "bonej_connectivity_density", "connectivity", "connectivity_density",
                    "roi_boundary_bone_fraction")] + [("geometry", BmGeometry)]


class BmBatchItem(ctypes.Structure):
    _fields_ = [("d_bone", ctypes.c_void_p), ("d_roi", ctypes.c_void_p),
                ("width", ctypes.c_uint64), ("height", ctypes.c_uint64),
                ("depth", ctypes.c_uint64), ("voxel_spacing", ctypes.c_double * 3),
                ("feature_mode", ctypes.c_int)]


class BmBatchResult(ctypes.Structure):
    _fields_ = [("metrics", BmThicknessSpacing), ("status", ctypes.c_int),
                ("message", ctypes.c_char * 252)]


class BmRequestTiming(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("start_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("batch_size", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class BmServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_uint64), ("requests_done", ctypes.c_uint64),
                ("max_batch", ctypes.c_uint64), ("sum_batch", ctypes.c_uint64),
                ("peak_device_bytes", ctypes.c_uint64), ("max_inflight", ctypes.c_uint32),
                ("slots", ctypes.c_uint32), ("busy_s", ctypes.c_double)]


class BmKernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint64),
                ("total_ms", ctypes.c_double)]


_lib = None
_synth = None


def lib():
    """The loaded libbm.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libbm.so is not built; run __graft_entry__.build() or "
                              "python bone_morphometry_amd/build.py")
        if os.path.exists(TUNE_PATH):
            os.environ.setdefault("BM_TUNE_CACHE", TUNE_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P, U64, I, VP = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p
        VOL = ctypes.POINTER(BmVolume)
        L.bm_last_error.restype = ctypes.c_char_p
        L.bm_last_error.argtypes = []
        L.bm_version.restype = I
        L.bm_device_count.argtypes = [ctypes.POINTER(I)]
        L.bm_binary_volume_check.argtypes = [VOL]
        L.bm_phase_inside.argtypes = [VOL, VOL, I, P, U64]
        L.bm_euclidean_distance_transform_squared.argtypes = [VOL, I, P, U64]
        L.bm_local_thickness.argtypes = [VOL, I, P, U64]
        L.bm_analyze_thickness_spacing.argtypes = [VOL, VOL, ctypes.POINTER(ctypes.c_double), I,
                                                   ctypes.POINTER(BmThicknessSpacing)]
        L.bm_euclidean_distance_transform_squared_device.argtypes = [P, U64, U64, U64, I, P, VP]
        L.bm_local_thickness_device.argtypes = [P, U64, U64, U64, I, P, VP]
        L.bm_analyze_thickness_spacing_device.argtypes = [
            P, P, U64, U64, U64, ctypes.POINTER(ctypes.c_double), I,
            ctypes.POINTER(BmThicknessSpacing), VP]
        L.bm_analyze_thickness_spacing_device_noalloc.argtypes = [
            P, P, U64, U64, U64, ctypes.POINTER(ctypes.c_double), I,
            ctypes.POINTER(BmThicknessSpacing), VP]
        L.bm_local_thickness_stages.argtypes = [VOL, I, P, P, P, P, ctypes.POINTER(U64),
                                                ctypes.POINTER(U64)]
        L.bm_set_stats_order.argtypes = [I, ctypes.POINTER(I)]
        ITEMS = ctypes.POINTER(BmBatchItem)
        L.bm_batch_create.argtypes = [ITEMS, U64, I, ctypes.POINTER(VP)]
        L.bm_batch_run.argtypes = [VP, ctypes.POINTER(BmBatchResult)]
        L.bm_batch_device_bytes.argtypes = [VP, ctypes.POINTER(U64)]
        L.bm_batch_destroy.argtypes = [VP]
        L.bm_autotune.argtypes = [ITEMS, U64, ctypes.c_char_p, I]
        L.bm_tuning_load.argtypes = [ctypes.c_char_p]
        L.bm_server_create.argtypes = [U64, I, I, ctypes.POINTER(VP)]
        L.bm_server_submit.argtypes = [VP, P, P, U64, U64, U64, ctypes.POINTER(ctypes.c_double),
                                       I, ctypes.POINTER(U64)]
        L.bm_server_poll.argtypes = [VP, U64, I, ctypes.POINTER(BmBatchResult),
                                     ctypes.POINTER(BmRequestTiming), ctypes.POINTER(I)]
        L.bm_server_cancel.argtypes = [VP, U64, ctypes.POINTER(I)]
        L.bm_server_stats_get.argtypes = [VP, ctypes.POINTER(BmServerStats)]
        L.bm_server_destroy.argtypes = [VP]
        D = ctypes.c_double
        L.bm_require_isotropic_spacing.argtypes = [ctypes.POINTER(D), ctypes.POINTER(D)]
        L.bm_gaussian_kernel.argtypes = [D, P, U64, ctypes.POINTER(U64)]
        L.bm_signed_distance_field_um.argtypes = [VOL, D, P, U64]
        L.bm_gaussian_smooth_nearest_in_place.argtypes = [P, U64, U64, U64, U64, D]
        L.bm_smoothed_signed_distance_field_um.argtypes = [VOL, D, D, P, U64]
        L.bm_signed_distance_field_um_device.argtypes = [P, U64, U64, U64, D, P, VP]
        L.bm_gaussian_smooth_nearest_in_place_device.argtypes = [P, U64, U64, U64, D, VP]
        L.bm_smoothed_signed_distance_field_um_device.argtypes = [P, U64, U64, U64, D, D, P, VP]
        L.bm_cubical_measures.argtypes = [VOL, ctypes.POINTER(BmCubical)]
        L.bm_analyze_topology.argtypes = [VOL, ctypes.c_int64, ctypes.POINTER(BmTopology)]
        L.bm_bonej_connectivity.argtypes = [VOL, ctypes.c_int64, ctypes.POINTER(BmBonej)]
        L.bm_bone_marrow_interface_faces.argtypes = [VOL, VOL, ctypes.POINTER(ctypes.c_int64)]
        L.bm_analyze.argtypes = [VOL, VOL, ctypes.POINTER(D), I, ctypes.POINTER(BmMetrics)]
        L.bm_analyze_device.argtypes = [P, P, U64, U64, U64, ctypes.POINTER(D), I,
                                        ctypes.POINTER(BmMetrics), VP]
        CS, U8 = ctypes.c_char_p, ctypes.c_uint8
        DIMS = ctypes.POINTER(U64)
        L.bm_read_binary_tiff.argtypes = [CS, U8, I, I, P, U64, DIMS]
        L.bm_read_roi_tiff.argtypes = [CS, I, P, U64, DIMS]
        L.bm_read_binary_tiff_device.argtypes = [CS, U8, I, I, P, U64, DIMS, VP]
        L.bm_read_roi_tiff_device.argtypes = [CS, I, P, U64, DIMS, VP]
        L.bm_read_isotropic_spacing_um.argtypes = [CS, ctypes.POINTER(D)]
        L.bm_profile_enable.argtypes = [I]
        L.bm_profile_read.argtypes = [ctypes.POINTER(BmKernelTime), I, ctypes.POINTER(I)]
        L.bm_profile_only.argtypes = [ctypes.c_char_p]
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("bm_last_error", "bm_version"):
                f.restype = I
        _lib = L
    return _lib


def synth():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise ImportError("libbm_synth.so is not built")
        S = ctypes.CDLL(SYNTH_PATH)
        P, Z, D = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double
        S.bm_synth_volume.restype = D
        S.bm_synth_volume.argtypes = [P, Z, Z, Z, D, D, D, D, ctypes.c_uint64, ctypes.c_int]
        S.bm_synth_roi_cylinder.restype = None
        S.bm_synth_roi_cylinder.argtypes = [P, Z, Z, Z, D, D, D]
        S.bm_synth_roi_ellipsoid.restype = None
        S.bm_synth_roi_ellipsoid.argtypes = [P, Z, Z, Z, D, D, D, D, D, D]
        _synth = S
    return _synth


def device_count():
    """HIP devices visible (0 when none or when libbm.so is missing)."""
    try:
        c = ctypes.c_int(0)
        lib().bm_device_count(ctypes.byref(c))
        return c.value
    except (ImportError, OSError):
        return 0


def profile_enable(on=True):
    lib().bm_profile_enable(1 if on else 0)


def profile_only(names=None):
    """Times only the kernels named (an iterable of names); None: every kernel."""
    lib().bm_profile_only(",".join(names).encode() if names else None)


def profile_read(max_kernels=64):
    """{kernel name: (launches, total_ms)} since the last read."""
    arr = (BmKernelTime * max_kernels)()
    n = ctypes.c_int(0)
    rc = lib().bm_profile_read(arr, max_kernels, ctypes.byref(n))
    if rc != BM_OK:
        raise RuntimeError(lib().bm_last_error().decode())
    return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}
