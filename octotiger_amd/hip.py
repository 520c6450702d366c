"""Minimal ctypes access to the HIP runtime that liboctofmm links
(/opt/rocm/lib/libamdhip64.so.7): device buffers, copies, streams and events
for callers of the level-batch API.  Using the library's own runtime keeps a
single HIP runtime per process.
"""
import ctypes
import os

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def rt():
    global _hip
    if _hip is None:
        from . import native
        native.lib()   # loads libamdhip64.so.7 as a dependency first
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        for name in ("libamdhip64.so.7", os.path.join(rocm, "lib", "libamdhip64.so.7"),
                     os.path.join(rocm, "lib", "libamdhip64.so")):
            try:
                _hip = ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("HIP runtime libamdhip64.so.7 not found")
        vp = ctypes.c_void_p
        _hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        _hip.hipFree.argtypes = [vp]
        _hip.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
        _hip.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipStreamDestroy.argtypes = [vp]
        _hip.hipStreamSynchronize.argtypes = [vp]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventDestroy.argtypes = [vp]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipGetErrorString.argtypes = [ctypes.c_int]
    return _hip


def check(err, what="HIP call"):
    if err != 0:
        raise RuntimeError(f"{what} failed: {rt().hipGetErrorString(err).decode()} ({err})")


def device_count():
    n = ctypes.c_int(0)
    err = rt().hipGetDeviceCount(ctypes.byref(n))
    return n.value if err == 0 else 0


def synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """float64 / int32 device buffer with the shape of a numpy array."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(rt().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(rt().hipMemcpy(d.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return d

    @property
    def addr(self):
        return self.ptr.value

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self):
        check(rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte buffer")
        check(rt().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.h = ctypes.c_void_p()
        check(rt().hipStreamCreate(ctypes.byref(self.h)), "hipStreamCreate")

    @property
    def handle(self):
        return self.h.value or 0

    def synchronize(self):
        check(rt().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        if self.h:
            rt().hipStreamDestroy(self.h)
            self.h = ctypes.c_void_p()


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        check(rt().hipEventCreate(ctypes.byref(self.h)), "hipEventCreate")

    def record(self, stream):
        check(rt().hipEventRecord(self.h, ctypes.c_void_p(stream.handle if isinstance(stream, Stream) else stream)),
              "hipEventRecord")

    def synchronize(self):
        check(rt().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, end):
        f = ctypes.c_float(0.0)
        check(rt().hipEventElapsedTime(ctypes.byref(f), self.h, end.h), "hipEventElapsedTime")
        return f.value

    def destroy(self):
        if self.h:
            rt().hipEventDestroy(self.h)
            self.h = ctypes.c_void_p()


class LevelOnDevice:
    """A workloads.Level's state uploaded to the GPU, with zeroed L / L_c."""

    def __init__(self, level):
        self.level = level
        self.M = DeviceArray.from_numpy(level.M)
        self.com = DeviceArray.from_numpy(level.com)
        self.mon = DeviceArray.from_numpy(level.mon)
        self.L = DeviceArray((level.nslots, 512, 20))
        self.L_c = DeviceArray((level.nslots, 512, 3))
        self.zero_outputs()

    def zero_outputs(self):
        self.L.zero()
        self.L_c.zero()

    def plan(self, ctx, **kw):
        return ctx.level_plan(self.level, self.M.addr, self.com.addr, self.mon.addr, self.L.addr,
                              self.L_c.addr, **kw)

    def free(self):
        for a in (self.M, self.com, self.mon, self.L, self.L_c):
            a.free()


class TreeOnDevice:
    """A tree.Tree on the GPU: per level the leaf masses (uploaded) and the
    arrays the solve writes (M, com, L, L_c, g)."""

    def __init__(self, tree):
        self.tree = tree
        self.levels = []
        for lv in tree.levels:
            n = lv.nnodes
            mon = lv.mon if lv.mon is not None else np.zeros((n, 512))
            d = {"M": DeviceArray((n, 512, 20)), "com": DeviceArray((n, 512, 3)),
                 "mon": DeviceArray.from_numpy(np.ascontiguousarray(mon, dtype=np.float64)),
                 "L": DeviceArray((n, 512, 20)), "L_c": DeviceArray((n, 512, 3)), "g": DeviceArray((n, 512, 4))}
            for k in ("M", "com", "L", "L_c", "g"):
                d[k].zero()
            self.levels.append(d)

    def addresses(self):
        return [{k: v.addr for k, v in d.items()} for d in self.levels]

    def plan(self, ctx, **kw):
        return ctx.tree_plan(self.tree, self.addresses(), **kw)

    def fetch(self, *keys):
        return [{k: d[k].numpy() for k in keys} for d in self.levels]

    def free(self):
        for d in self.levels:
            for a in d.values():
                a.free()
