"""ctypes binding of oracle/liboracle.so (the CPU restatement; tests only)."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")

Z, P, I, D = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int, ctypes.c_double


class LtDebug(ctypes.Structure):
    _fields_ = [("ridge_flags", P), ("propagated", P), ("squared_edt", P),
                ("ridge_count", ctypes.c_uint64), ("fill_visits", ctypes.c_uint64),
                ("propagate_from_all", I), ("reverse_cleanup", I)]


class Stats(ctypes.Structure):
    _fields_ = [(n, D) for n in ("mean", "standard_deviation", "maximum", "median",
                                 "percentile_10", "percentile_25", "percentile_75",
                                 "percentile_90", "interquartile_range",
                                 "coefficient_of_variation")] + [("count", ctypes.c_uint64)]


class Metrics(ctypes.Structure):
    _fields_ = [("thickness", Stats), ("spacing", Stats), ("bone_volume", D),
                ("total_volume", D), ("bone_volume_fraction", D)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(PATH)
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_phase_inside.argtypes = [P, P, Z, I, P]
        L.oracle_phase_inside.restype = None
        L.oracle_edt_squared.argtypes = [P, Z, Z, Z, I, P]
        L.oracle_templates.argtypes = [P, Z, P]
        L.oracle_templates.restype = None
        L.oracle_local_thickness_ex.argtypes = [P, Z, Z, Z, I, P, ctypes.POINTER(LtDebug)]
        L.oracle_local_thickness.argtypes = [P, Z, Z, Z, I, P]
        L.oracle_nonzero_stats.argtypes = [P, Z, D, I, ctypes.POINTER(Stats)]
        L.oracle_quantile.argtypes = [P, Z, D, ctypes.c_float]
        L.oracle_quantile.restype = D
        L.oracle_analyze_thickness_spacing.argtypes = [P, P, Z, Z, Z, ctypes.POINTER(D), I, I,
                                                       ctypes.POINTER(Metrics)]
        L.oracle_set_num_threads.argtypes = [I]
        L.oracle_set_num_threads.restype = None
        L.oracle_num_threads.restype = I
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _chk(rc):
    if rc != 0:
        raise OracleError(lib().oracle_last_error().decode())


def phase_inside(bone, roi, select_bone):
    bone = np.ascontiguousarray(bone, np.uint8)
    roi = np.ascontiguousarray(roi, np.uint8)
    out = np.empty(bone.size, np.uint8)
    lib().oracle_phase_inside(bone.ctypes.data, roi.ctypes.data, bone.size, int(select_bone),
                              out.ctypes.data)
    return out


def edt_squared(vol, w, h, d, phase=0):
    vol = np.ascontiguousarray(vol, np.uint8).reshape(-1)
    out = np.empty(vol.size, np.uint32)
    _chk(lib().oracle_edt_squared(vol.ctypes.data, w, h, d, int(phase), out.ctypes.data))
    return out


def templates(values):
    v = np.ascontiguousarray(values, np.uint32)
    out = np.empty(3 * v.size, np.uint32)
    lib().oracle_templates(v.ctypes.data, v.size, out.ctypes.data)
    return out.reshape(3, v.size)


def local_thickness_stages(vol, w, h, d, phase=0, propagate_from_all=False,
                           reverse_cleanup=False):
    vol = np.ascontiguousarray(vol, np.uint8).reshape(-1)
    n = vol.size
    out = np.empty(n, np.float32)
    flags = np.empty(n, np.uint8)
    prop = np.empty(n, np.uint32)
    sq = np.empty(n, np.uint32)
    dbg = LtDebug(flags.ctypes.data, prop.ctypes.data, sq.ctypes.data, 0, 0,
                  int(propagate_from_all), int(reverse_cleanup))
    _chk(lib().oracle_local_thickness_ex(vol.ctypes.data, w, h, d, int(phase), out.ctypes.data,
                                         ctypes.byref(dbg)))
    return dict(thickness=out, ridge_flags=flags, propagated=prop, squared_edt=sq,
                ridge_count=dbg.ridge_count, fill_visits=dbg.fill_visits)


def local_thickness(vol, w, h, d, phase=0):
    return local_thickness_stages(vol, w, h, d, phase)["thickness"]


def nonzero_stats(values, scale, include_quantiles=True):
    v = np.ascontiguousarray(values, np.float32)
    st = Stats()
    _chk(lib().oracle_nonzero_stats(v.ctypes.data, v.size, float(scale), int(include_quantiles),
                                    ctypes.byref(st)))
    return st


def quantile(values, q, scale):
    v = np.ascontiguousarray(values, np.float32)
    return lib().oracle_quantile(v.ctypes.data, v.size, float(q), float(scale))


def analyze_thickness_spacing(bone, roi, w, h, d, spacing, bonej_mode=False,
                              include_quantiles=True):
    bone = np.ascontiguousarray(bone, np.uint8).reshape(-1)
    roi = np.ascontiguousarray(roi, np.uint8).reshape(-1)
    sp = (D * 3)(*spacing)
    m = Metrics()
    _chk(lib().oracle_analyze_thickness_spacing(bone.ctypes.data, roi.ctypes.data, w, h, d, sp,
                                                int(bonej_mode), int(include_quantiles),
                                                ctypes.byref(m)))
    return m


def set_threads(n):
    lib().oracle_set_num_threads(int(n))


def num_threads():
    return lib().oracle_num_threads()


def _bind_smoothing():
    L = lib()
    if getattr(L, "_smoothing_bound", False):
        return L
    L.oracle_gaussian_kernel.argtypes = [D, P, Z]
    L.oracle_gaussian_kernel.restype = Z
    L.oracle_gaussian_smooth_nearest.argtypes = [P, Z, Z, Z, D]
    L.oracle_signed_distance_field_um.argtypes = [P, Z, Z, Z, D, P]
    L._smoothing_bound = True
    return L


def gaussian_kernel(sigma):
    """gaussian_kernel, src/sdf_curvature.rs:182-198."""
    L = _bind_smoothing()
    n = L.oracle_gaussian_kernel(float(sigma), None, 0)
    if n == 0:
        raise OracleError(L.oracle_last_error().decode())
    out = np.empty(n, np.float64)
    L.oracle_gaussian_kernel(float(sigma), out.ctypes.data, n)
    return out


def gaussian_smooth_nearest(data, w, h, d, sigma_vox):
    """gaussian_smooth_nearest_in_place, src/sdf_curvature.rs:200-269, on a copy."""
    L = _bind_smoothing()
    out = np.array(data, dtype=np.float64, copy=True).reshape(-1)
    _chk(L.oracle_gaussian_smooth_nearest(out.ctypes.data, w, h, d, float(sigma_vox)))
    return out


def signed_distance_field_um(bone, w, h, d, hs):
    """signed_distance_field_um, src/sdf_curvature.rs:162-180."""
    L = _bind_smoothing()
    bone = np.ascontiguousarray(bone, np.uint8).reshape(-1)
    out = np.empty(bone.size, np.float64)
    _chk(L.oracle_signed_distance_field_um(bone.ctypes.data, w, h, d, float(hs),
                                           out.ctypes.data))
    return out


class Cubical(ctypes.Structure):
    _fields_ = [("raw_euler", ctypes.c_int64), ("mean_breadth_lattice_units", D),
                ("vertices", ctypes.c_int64), ("edges", ctypes.c_int64),
                ("faces", ctypes.c_int64), ("cubes", ctypes.c_int64)]


class Bonej(ctypes.Structure):
    _fields_ = [("edge_correction", D), ("delta_chi", D), ("connectivity", D)] + [
        (n, ctypes.c_int64) for n in ("f", "e", "c", "d", "a", "b")]


class Topology(ctypes.Structure):
    _fields_ = [("raw_euler", ctypes.c_int64), ("beta0", ctypes.c_uint64),
                ("beta1", ctypes.c_uint64), ("beta2", ctypes.c_uint64)]


class Rest(ctypes.Structure):
    _fields_ = [("bone_volume", D), ("total_volume", D), ("bone_volume_fraction", D),
                ("topology", Topology), ("bonej", Bonej),
                ("bonej_connectivity_density", D), ("connectivity", D),
                ("connectivity_density", D), ("roi_boundary_bone_fraction", D),
                ("bone_surface", D), ("bone_surface_to_bone_volume", D),
                ("bone_surface_to_total_volume", D), ("mean_breadth", D),
                ("interface_faces", ctypes.c_int64), ("cubical", Cubical)]


def _bind_topology():
    L = lib()
    if getattr(L, "_topology_bound", False):
        return L
    L.oracle_cubical_measures.argtypes = [P, Z, Z, Z, ctypes.POINTER(Cubical)]
    L.oracle_cubical_measures.restype = None
    L.oracle_bonej_connectivity.argtypes = [P, Z, Z, Z, ctypes.c_int64, ctypes.POINTER(Bonej)]
    L.oracle_bonej_connectivity.restype = None
    L.oracle_analyze_topology.argtypes = [P, Z, Z, Z, ctypes.c_int64, ctypes.POINTER(Topology)]
    L.oracle_interface_faces.argtypes = [P, P, Z, Z, Z]
    L.oracle_interface_faces.restype = ctypes.c_int64
    L.oracle_roi_boundary_counts.argtypes = [P, P, Z, Z, Z, ctypes.POINTER(ctypes.c_uint64),
                                             ctypes.POINTER(ctypes.c_uint64)]
    L.oracle_roi_boundary_counts.restype = None
    L.oracle_analyze_rest.argtypes = [P, P, Z, Z, Z, ctypes.POINTER(D), ctypes.POINTER(Rest)]
    L._topology_bound = True
    return L


def _u8(a):
    return np.ascontiguousarray(a, np.uint8).reshape(-1)


def cubical_measures(vol, w, h, d):
    """cubical_measures, src/topology.rs:53-171."""
    L = _bind_topology()
    vol = _u8(vol)
    out = Cubical()
    L.oracle_cubical_measures(vol.ctypes.data, w, h, d, ctypes.byref(out))
    return out


def bonej_connectivity(vol, w, h, d, raw_euler):
    """bonej_connectivity, src/topology.rs:175-201."""
    L = _bind_topology()
    vol = _u8(vol)
    out = Bonej()
    L.oracle_bonej_connectivity(vol.ctypes.data, w, h, d, int(raw_euler), ctypes.byref(out))
    return out


def analyze_topology(vol, w, h, d, raw_euler):
    """analyze_topology, src/topology.rs:29-51."""
    L = _bind_topology()
    vol = _u8(vol)
    out = Topology()
    _chk(L.oracle_analyze_topology(vol.ctypes.data, w, h, d, int(raw_euler), ctypes.byref(out)))
    return out


def interface_faces(bone, roi, w, h, d):
    """bone_marrow_interface_faces, src/topology.rs:442-490."""
    L = _bind_topology()
    bone, roi = _u8(bone), _u8(roi)
    return L.oracle_interface_faces(bone.ctypes.data, roi.ctypes.data, w, h, d)


def roi_boundary_counts(abone, roi, w, h, d):
    """The two counts of roi_boundary_bone_fraction, src/metrics.rs:166-202."""
    L = _bind_topology()
    abone, roi = _u8(abone), _u8(roi)
    b, t = ctypes.c_uint64(), ctypes.c_uint64()
    L.oracle_roi_boundary_counts(abone.ctypes.data, roi.ctypes.data, w, h, d, ctypes.byref(b),
                                 ctypes.byref(t))
    return b.value, t.value


def analyze_rest(bone, roi, w, h, d, spacing):
    """The Metrics fields of analyze beyond thickness/spacing, src/metrics.rs:123-164."""
    L = _bind_topology()
    bone, roi = _u8(bone), _u8(roi)
    sp = (D * 3)(*[float(s) for s in spacing])
    out = Rest()
    _chk(L.oracle_analyze_rest(bone.ctypes.data, roi.ctypes.data, w, h, d, sp, ctypes.byref(out)))
    return out
