"""Device buffers and events on the HIP runtime libpado_gpu.so is linked with.

torch's wheel bundles its own HIP and HSA runtimes; two runtimes cannot both
drive the GPU from one process, so GPU-side plumbing in this package (bench,
tests) goes through the runtime the library itself loaded, found by path in
the process map after the library is loaded.
"""
import ctypes

import numpy as np

_rt = None


def runtime():
    """ctypes handle of the libamdhip64 that libpado_gpu.so uses."""
    global _rt
    if _rt is not None:
        return _rt
    from . import lib

    lib()  # load libpado_gpu.so (and with it its HIP runtime)
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so.7" in line:
                path = line.split()[-1]
                break
    if path is None:
        raise RuntimeError("the HIP runtime of libpado_gpu.so is not mapped")
    rt = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    rt.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    rt.hipFree.argtypes = [vp]
    rt.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    rt.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
    rt.hipDeviceSynchronize.argtypes = []
    rt.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    rt.hipEventRecord.argtypes = [vp, vp]
    rt.hipEventSynchronize.argtypes = [vp]
    rt.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    rt.hipEventDestroy.argtypes = [vp]
    _rt = rt
    return rt


def _ok(code, what):
    if code != 0:
        raise RuntimeError(f"{what} failed with HIP error {code}")


class DeviceArray:
    """A flat device buffer holding a numpy dtype."""

    H2D, D2H = 1, 2

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(self.count * self.dtype.itemsize, 1)
        p = ctypes.c_void_p()
        _ok(runtime().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            _ok(runtime().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, cls.H2D), "hipMemcpy H2D")
        return d

    def numpy(self):
        out = np.empty(self.count, self.dtype)
        if self.count:
            _ok(runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        _ok(runtime().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Events:
    """hipEvent timing on an explicit stream (e.g. the context's stream)."""

    def __init__(self):
        self.rt = runtime()
        self._all = []

    def event(self):
        e = ctypes.c_void_p()
        _ok(self.rt.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self._all.append(e)
        return e

    def record(self, e, stream):
        _ok(self.rt.hipEventRecord(e, ctypes.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, a, b):
        _ok(self.rt.hipEventSynchronize(b), "hipEventSynchronize")
        t = ctypes.c_float()
        _ok(self.rt.hipEventElapsedTime(ctypes.byref(t), a, b), "hipEventElapsedTime")
        return float(t.value)

    def close(self):
        for e in self._all:
            self.rt.hipEventDestroy(e)
        self._all = []
