"""Minimal ctypes access to the HIP runtime: copies, synchronisation, streams, events.

Used by the Python host mirror to move state in and out of the engine's device
buffers and to time kernels on the engine's own stream."""
import ctypes as C
import os

import numpy as np

_hip = None


def lib():
    global _hip
    if _hip is None:
        for p in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(p)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        h = _hip
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemcpy.restype = C.c_int
        h.hipDeviceSynchronize.restype = C.c_int
        h.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        h.hipGetDeviceCount.restype = C.c_int
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamCreate.restype = C.c_int
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.restype = C.c_int
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventCreate.restype = C.c_int
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventRecord.restype = C.c_int
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventElapsedTime.restype = C.c_int
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipMalloc.restype = C.c_int
        h.hipFree.argtypes = [C.c_void_p]
        h.hipFree.restype = C.c_int
        h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        h.hipMemcpyAsync.restype = C.c_int
        h.hipGetErrorString.argtypes = [C.c_int]
        h.hipGetErrorString.restype = C.c_char_p
    return _hip


H2D, D2H, D2D = 1, 2, 3


def _chk(code, what):
    if code != 0:
        raise RuntimeError(f"{what}: {lib().hipGetErrorString(code).decode()}")


def device_count():
    n = C.c_int(0)
    if lib().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def to_device(dev_ptr, arr):
    _chk(lib().hipMemcpy(dev_ptr, arr.ctypes.data, arr.nbytes, H2D), "hipMemcpy H2D")


def from_device(arr, dev_ptr):
    _chk(lib().hipMemcpy(arr.ctypes.data, dev_ptr, arr.nbytes, D2H), "hipMemcpy D2H")


def copy_d2d(dst_ptr, src_ptr, nbytes, stream=None):
    _chk(lib().hipMemcpyAsync(dst_ptr, src_ptr, nbytes, D2D, stream), "hipMemcpyAsync D2D")


class DeviceArray:
    """A float64 device buffer owned by the harness (hipMalloc / hipFree)."""

    def __init__(self, n):
        self.n = int(n)
        p = C.c_void_p()
        _chk(lib().hipMalloc(C.byref(p), max(self.n, 1) * 8), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        d = cls(a.size)
        if a.size:
            to_device(d.ptr, a)
        return d

    def get(self):
        out = np.empty(self.n)
        synchronize()
        if self.n:
            from_device(out, self.ptr)
        return out

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def synchronize():
    _chk(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _chk(lib().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s

    def synchronize(self):
        _chk(lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")


class Graph:
    """Stream capture into a hipGraph: `with g.capture(stream): ...` records the work issued
    on `stream`; launch() replays it."""

    def __init__(self):
        h = lib()
        h.hipStreamBeginCapture.argtypes = [C.c_void_p, C.c_int]
        h.hipStreamBeginCapture.restype = C.c_int
        h.hipStreamEndCapture.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        h.hipStreamEndCapture.restype = C.c_int
        h.hipGraphInstantiate.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t]
        h.hipGraphInstantiate.restype = C.c_int
        h.hipGraphLaunch.argtypes = [C.c_void_p, C.c_void_p]
        h.hipGraphLaunch.restype = C.c_int
        h.hipGraphExecDestroy.argtypes = [C.c_void_p]
        h.hipGraphDestroy.argtypes = [C.c_void_p]
        self.graph, self.exec = C.c_void_p(), C.c_void_p()

    class _Capture:
        def __init__(self, g, stream):
            self.g, self.stream = g, stream

        def __enter__(self):
            # relaxed mode: a host-side free by the garbage collector does not void the capture
            _chk(lib().hipStreamBeginCapture(self.stream.handle, 2), "hipStreamBeginCapture")
            return self.g

        def __exit__(self, et, ev, tb):
            rc = lib().hipStreamEndCapture(self.stream.handle, C.byref(self.g.graph))
            if et is None:
                _chk(rc, "hipStreamEndCapture")
                _chk(lib().hipGraphInstantiate(C.byref(self.g.exec), self.g.graph, None, None, 0), "hipGraphInstantiate")
            return False

    def capture(self, stream):
        return Graph._Capture(self, stream)

    def launch(self, stream):
        _chk(lib().hipGraphLaunch(self.exec, stream.handle), "hipGraphLaunch")

    def __del__(self):
        try:
            if self.exec:
                lib().hipGraphExecDestroy(self.exec)
            if self.graph:
                lib().hipGraphDestroy(self.graph)
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        _chk(lib().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e

    def record(self, stream=None):
        _chk(lib().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def elapsed_ms(self, end):
        lib().hipEventSynchronize(end.handle)
        ms = C.c_float()
        _chk(lib().hipEventElapsedTime(C.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return ms.value
