"""Helpers for GPU tests: device buffers through libfgx's own allocator and copies.

Everything runs on the one HIP runtime libfgx links (torch ships a second one;
two runtimes in one process can leave the device to whichever starts first).
"""
import ctypes

import numpy as np

from flamegpu2_amd import fgx

_ctx = None


def ctx():
    global _ctx
    if _ctx is None:
        _ctx = fgx.Context(0)
    return _ctx


class DeviceArray:
    """A device allocation holding `n` elements of `dtype`."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = max(int(n), 1)
        self.nbytes = self.n * self.dtype.itemsize
        self.alloc_bytes = (self.nbytes + 3) // 4 * 4
        p = ctypes.c_void_p()
        fgx.check(fgx.lib().fgx_malloc(ctx().h, self.alloc_bytes, ctypes.byref(p)))
        self.addr = p.value

    def data_ptr(self):
        return self.addr

    def numel(self):
        return self.n

    def __del__(self):
        try:
            if self.addr:
                fgx.lib().fgx_free(ctx().h, ctypes.c_void_p(self.addr))
                self.addr = None
        except Exception:
            pass


def dev(a):
    a = np.ascontiguousarray(a)
    d = DeviceArray(len(a), a.dtype)
    if a.size:
        fgx.check(fgx.lib().fgx_memcpy_h2d(ctypes.c_void_p(d.addr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, None))
    sync()
    return d


def empty(n, dtype):
    d = DeviceArray(n, dtype)
    fgx.check(fgx.lib().fgx_fill_u32(ctypes.c_void_p(d.addr), 0, d.alloc_bytes // 4, None))
    sync()
    return d


def host(t, dtype=None, n=None):
    sync()
    a = np.empty(t.n, dtype=t.dtype)
    fgx.check(fgx.lib().fgx_memcpy_d2h(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(t.addr), t.nbytes, None))
    sync()
    if dtype is not None:
        a = a.view(dtype)
    if n is not None:
        a = a[:n]
    return a


def ptr(t):
    return t.addr


def sync():
    fgx.check(fgx.lib().fgx_device_synchronize())


def scratch(nbytes):
    s = DeviceArray(max(nbytes // 4, 64), np.uint32)
    fgx.scratch_clear(s.addr, s.nbytes)
    sync()
    return s
