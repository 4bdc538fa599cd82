"""Minimal ctypes access to the HIP runtime that libmmc.so links
(/opt/rocm/lib/libamdhip64.so.7), for tests that hand device memory and
streams to the C-ABI.  torch is not used for this: its wheel carries its own
HIP runtime, and two runtimes in one process do not share devices."""
import ctypes as C

import numpy as np

_rt = None
MEMCPY_H2D, MEMCPY_D2H = 1, 2
CAPTURE_GLOBAL = 0


def rt():
    global _rt
    if _rt is None:
        L = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        P, PP = C.c_void_p, C.POINTER(C.c_void_p)
        sig = {
            "hipMalloc": [PP, C.c_size_t], "hipFree": [P],
            "hipMemcpy": [P, P, C.c_size_t, C.c_int],
            "hipDeviceSynchronize": [], "hipStreamCreate": [PP], "hipStreamDestroy": [P],
            "hipStreamSynchronize": [P], "hipStreamBeginCapture": [P, C.c_int],
            "hipStreamEndCapture": [P, PP], "hipGraphInstantiateWithFlags": [PP, P, C.c_ulonglong],
            "hipGraphLaunch": [P, P], "hipGraphExecDestroy": [P], "hipGraphDestroy": [P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes, fn.restype = args, C.c_int
        _rt = L
    return _rt


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device buffer holding a float64 numpy array (freed on close)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float64)
        self.n = host.size
        self.ptr = C.c_void_p()
        check(rt().hipMalloc(C.byref(self.ptr), max(8, host.nbytes)), "hipMalloc")
        self.write(host)

    def write(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.size == self.n
        check(rt().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, MEMCPY_H2D), "hipMemcpy H2D")

    def read(self) -> np.ndarray:
        out = np.empty(self.n)
        check(rt().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, MEMCPY_D2H), "hipMemcpy D2H")
        return out

    def close(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    __del__ = close


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        check(rt().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def sync(self):
        check(rt().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def close(self):
        if self.h:
            rt().hipStreamDestroy(self.h)
            self.h = C.c_void_p()

    __del__ = close


class Graph:
    """Captures the work queued on `stream` by fn() (global capture mode: an
    allocation during capture is an error) into an executable graph."""

    def __init__(self, stream: Stream, fn):
        self.stream = stream
        check(rt().hipStreamBeginCapture(stream.h, CAPTURE_GLOBAL), "hipStreamBeginCapture")
        try:
            fn()
        finally:
            self.graph = C.c_void_p()
            check(rt().hipStreamEndCapture(stream.h, C.byref(self.graph)), "hipStreamEndCapture")
        self.exec = C.c_void_p()
        check(rt().hipGraphInstantiateWithFlags(C.byref(self.exec), self.graph, 0), "hipGraphInstantiate")

    def replay(self):
        check(rt().hipGraphLaunch(self.exec, self.stream.h), "hipGraphLaunch")

    def close(self):
        if getattr(self, "exec", None):
            rt().hipGraphExecDestroy(self.exec)
            self.exec = None
        if getattr(self, "graph", None):
            rt().hipGraphDestroy(self.graph)
            self.graph = None

    __del__ = close
