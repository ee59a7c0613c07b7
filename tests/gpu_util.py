"""Helpers for GPU tests: torch provides device memory; the library does the work."""
import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        assert t.cuda.is_available(), "GPU tests need a GPU"
        _torch = t
    return _torch


def dev(a):
    t = torch()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return t.from_numpy(a.view(np.int32)).cuda()
    if a.dtype == np.uint64:
        return t.from_numpy(a.view(np.int64)).cuda()
    return t.from_numpy(a).cuda()


def empty(n, dtype):
    t = torch()
    tmap = {np.float32: t.float32, np.float64: t.float64, np.uint32: t.int32, np.int32: t.int32,
            np.uint8: t.uint8, np.uint64: t.int64, np.int64: t.int64}
    return t.zeros(max(n, 1), dtype=tmap[np.dtype(dtype).type], device="cuda")


def host(t, dtype=None, n=None):
    a = t.detach().cpu().numpy()
    if dtype is not None:
        a = a.view(dtype)
    if n is not None:
        a = a[:n]
    return a


def ptr(t):
    return t.data_ptr()


def sync():
    torch().cuda.synchronize()


def scratch(nbytes):
    from flamegpu2_amd import fgx
    s = torch().zeros(max(nbytes // 4, 64), dtype=torch().int32, device="cuda")
    fgx.scratch_clear(s.data_ptr(), s.numel() * 4)
    return s
