"""Minimal ctypes access to the HIP runtime for device-resident buffers.

Used by callers that keep queries and results on the GPU (bfg_matrix_device)
and by the benchmark; the engine itself manages its own memory.
"""
import ctypes as C
import os

import numpy as np

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [C.c_int]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _hip.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        _hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        _hip.hipHostFree.argtypes = [C.c_void_p]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipSetDevice",
                  "hipMemGetInfo", "hipGetDeviceCount", "hipHostMalloc", "hipHostFree"):
            getattr(_hip, f).restype = C.c_int
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


def device_count():
    n = C.c_int(0)
    if hip().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def sync():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    _ok(hip().hipMemGetInfo(C.byref(f), C.byref(t)), "hipMemGetInfo")
    return f.value, t.value


class DeviceArray:
    """A device allocation of `count` elements of numpy dtype `dtype`."""

    def __init__(self, count, dtype=np.uint32):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(self.count * self.dtype.itemsize, 1)
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p

    @staticmethod
    def from_host(a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.size, a.dtype)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def to_host(self, start=0, count=None):
        count = self.count - start if count is None else count
        out = np.empty(count, dtype=self.dtype)
        src = C.c_void_p(self.ptr.value + start * self.dtype.itemsize)
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), src, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory (hipHostMalloc) of `count` elements, seen as a
    numpy array `a`: device-to-host copies into it run at the link's DMA rate
    instead of going through the runtime's staging buffers."""

    def __init__(self, count, dtype=np.uint32):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        nbytes = max(self.count * self.dtype.itemsize, 1)
        p = C.c_void_p()
        _ok(hip().hipHostMalloc(C.byref(p), nbytes, 0), "hipHostMalloc")
        self.ptr = p
        buf = (C.c_uint8 * nbytes).from_address(p.value)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=self.count)

    def free(self):
        if self.ptr is not None and self.ptr.value:
            self.a = None
            hip().hipHostFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
