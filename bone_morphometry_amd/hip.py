"""Minimal ctypes access to the HIP runtime: device buffers, streams, events.

Plumbing for the benchmark and the device-resident entry points (a torch
tensor's data_ptr() works as well); no compute happens here.
"""
import ctypes
import ctypes.util
import os

_hip = None


def rt():
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so",
                     ctypes.util.find_library("amdhip64")):
            if not cand:
                continue
            try:
                _hip = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise ImportError("libamdhip64.so not found")
        H = _hip
        P, Z = ctypes.c_void_p, ctypes.c_size_t
        H.hipMalloc.argtypes = [ctypes.POINTER(P), Z]
        H.hipFree.argtypes = [P]
        H.hipMemcpy.argtypes = [P, P, Z, ctypes.c_int]
        H.hipMemset.argtypes = [P, ctypes.c_int, Z]
        H.hipDeviceSynchronize.argtypes = []
        H.hipStreamCreate.argtypes = [ctypes.POINTER(P)]
        H.hipStreamSynchronize.argtypes = [P]
        H.hipEventCreate.argtypes = [ctypes.POINTER(P)]
        H.hipEventRecord.argtypes = [P, P]
        H.hipEventSynchronize.argtypes = [P]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), P, P]
        H.hipGetErrorString.restype = ctypes.c_char_p
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        H.hipMemGetInfo.argtypes = [ctypes.POINTER(Z), ctypes.POINTER(Z)]
    return _hip


H2D, D2H, D2D = 1, 2, 3


def check(rc):
    if rc != 0:
        raise RuntimeError("HIP error %d: %s" % (rc, rt().hipGetErrorString(rc).decode()))


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(rt().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)))

    @property
    def value(self):
        return self.ptr.value

    def upload(self, arr):
        assert arr.nbytes <= self.nbytes
        check(rt().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, H2D))

    def download(self, arr):
        assert arr.nbytes <= self.nbytes
        check(rt().hipMemcpy(arr.ctypes.data, self.ptr, arr.nbytes, D2H))

    def free(self):
        if self.ptr:
            check(rt().hipFree(self.ptr))
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def stream():
    s = ctypes.c_void_p()
    check(rt().hipStreamCreate(ctypes.byref(s)))
    return s


class Event:
    def __init__(self):
        self.e = ctypes.c_void_p()
        check(rt().hipEventCreate(ctypes.byref(self.e)))

    def record(self, s=None):
        check(rt().hipEventRecord(self.e, s))

    def elapsed_ms(self, end):
        check(rt().hipEventSynchronize(end.e))
        ms = ctypes.c_float()
        check(rt().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e))
        return ms.value


def synchronize():
    check(rt().hipDeviceSynchronize())


def mem_info():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    check(rt().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)))
    return free.value, total.value
