"""Device-memory plumbing for the GPU tests through the library's own C ABI
(ntc_dev_*); every kernel call passes the null stream."""
from __future__ import annotations

import ctypes as C

import numpy as np

import openntc_amd
from openntc_amd import _lib as L


# Every buffer made here stays alive until release() (called after each test
# by the autouse fixture in conftest.py, after a device synchronisation), so a
# call like f(p(dev(x))) never hands a kernel a pointer that Python has
# already freed.
_live: list = []


def release():
    if _live:
        lib().ntc_dev_synchronize()
        _live.clear()


class DevArray:
    """A device buffer with a numpy shape/dtype, freed by release()."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ok(lib().ntc_dev_alloc(self.nbytes, C.byref(p)), "ntc_dev_alloc")
        self.ptr = p.value
        _live.append(self)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().ntc_dev_free(self.ptr)
            self.ptr = None

    def data_ptr(self) -> int:
        return self.ptr


def lib():
    return openntc_amd.load()


def ok(status, what="call"):
    L.check(status, what)


def dev(a: np.ndarray) -> DevArray:
    a = np.ascontiguousarray(a)
    d = DevArray(a.shape, a.dtype)
    ok(lib().ntc_dev_upload(d.ptr, a.ctypes.data, a.nbytes), "upload")
    return d


def empty(shape, dtype) -> DevArray:
    return DevArray(shape, dtype)


def host(d: DevArray, dtype=None) -> np.ndarray:
    a = np.empty(d.shape, d.dtype)
    ok(lib().ntc_dev_download(a.ctypes.data, d.ptr, d.nbytes), "download")
    return a.view(dtype) if dtype is not None else a


def p(d: DevArray) -> int:
    return d.ptr


def feature_desc(fd: dict) -> L.FeatureDesc:
    return L.FeatureDesc(fd["g0_dim"], fd["g1_dim"], fd["c0"], fd["c1"], fd["g0_mode"],
                         fd["pe_first"], fd["pe_count"], fd["lod_slot"], fd["F"])


def tiles(origins_dev, n_tiles, T, M, norm_lod) -> L.Tiles:
    return L.Tiles(origins_dev.data_ptr(), n_tiles, T, M, norm_lod)


def byref(x):
    return C.byref(x)
