"""ctypes binding of the serial CPU oracle (oracle/liboracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the library itself.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
f32p = ctypes.POINTER(ctypes.c_float)


class SpatialMeta(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32), ("min", ctypes.c_float * 3), ("max", ctypes.c_float * 3),
                ("radius", ctypes.c_float), ("width", ctypes.c_float * 3), ("grid", ctypes.c_uint32 * 3),
                ("bins", ctypes.c_uint32), ("wrap_ok", ctypes.c_int32), ("cell", ctypes.c_float),
                ("span", ctypes.c_int32)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(_LIB)
        L.orc_rng_u32.restype = ctypes.c_uint32
        L.orc_rng_u32.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_rng_uniform.restype = ctypes.c_float
        L.orc_rng_uniform.argtypes = L.orc_rng_u32.argtypes
        L.orc_spatial_hash.restype = ctypes.c_uint32
        L.orc_spatial_hash.argtypes = [ctypes.POINTER(SpatialMeta), ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_spatial_candidates.restype = ctypes.c_uint32
        L.orc_spatial_candidates.argtypes = [ctypes.POINTER(SpatialMeta), u32p, ctypes.c_int, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_float, u32p, ctypes.c_uint32]
        L.orc_spatial_mean_candidates.restype = ctypes.c_double
        L.orc_spatial_mean_candidates.argtypes = [ctypes.POINTER(SpatialMeta), u32p, ctypes.c_int]
        L.orc_compact.restype = ctypes.c_uint32
        L.orc_births.restype = ctypes.c_uint32
        L.orc_sim_create.restype = ctypes.c_void_p
        L.orc_sim_create.argtypes = [ctypes.c_int, f32p, ctypes.c_uint32, ctypes.c_uint64]
        L.orc_sim_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sim_var_count.restype = ctypes.c_uint32
        L.orc_sim_var_count.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_var_bytes.restype = ctypes.c_uint32
        L.orc_sim_var_bytes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32]
        L.orc_sim_set_agents.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_void_p), u32p]
        L.orc_sim_get_agents.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), u32p]
        L.orc_sim_count.restype = ctypes.c_uint32
        L.orc_sim_count.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_step.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.orc_sim_bins.restype = ctypes.c_uint32
        L.orc_sim_bins.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_sim_get_pbm.argtypes = [ctypes.c_void_p, ctypes.c_int, u32p]
        L.orc_sim_set_bin_mult.argtypes = [ctypes.c_void_p, ctypes.c_float]
        L.orc_sim_last_candidates.restype = ctypes.c_double
        L.orc_sim_last_candidates.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def ptr(a, t=ctypes.c_uint32):
    return a.ctypes.data_as(ctypes.POINTER(t))


def philox(ctr, key):
    C = (ctypes.c_uint32 * 4)(*ctr)
    K = (ctypes.c_uint32 * 2)(*key)
    O = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(C, K, O)
    return list(O)


def spatial_meta(dims, mn, mx, radius, mult=1.0):
    m = SpatialMeta()
    mn = (ctypes.c_float * 3)(*(list(mn) + [0.0] * (3 - len(mn))))
    mx = (ctypes.c_float * 3)(*(list(mx) + [0.0] * (3 - len(mx))))
    rc = lib().orc_spatial_meta_init_mult(dims, mn, mx, ctypes.c_float(radius), ctypes.c_float(mult), ctypes.byref(m))
    assert rc == 0
    return m


def spatial_build(meta, x, y, z=None, extra=()):
    """Returns (pbm, perm, sorted arrays [x, y, (z), *extra])."""
    x = _f32(x); y = _f32(y)
    n = len(x)
    arrays = [x, y] + ([_f32(z)] if z is not None else []) + [np.ascontiguousarray(e) for e in extra]
    outs = [np.empty_like(a) for a in arrays]
    ins = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    outp = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in outs])
    nb = (ctypes.c_uint32 * len(arrays))(*[a.itemsize for a in arrays])
    pbm = np.zeros(meta.bins + 1, dtype=np.uint32)
    perm = np.zeros(n, dtype=np.uint32)
    zz = ptr(_f32(z), ctypes.c_float) if z is not None else None
    if z is not None:
        zarr = _f32(z)
        zz = ptr(zarr, ctypes.c_float)
    lib().orc_spatial_build(ctypes.byref(meta), ctypes.c_uint32(n), ptr(x, ctypes.c_float), ptr(y, ctypes.c_float),
                            zz, ctypes.c_uint32(len(arrays)), ins, nb, outp, ptr(pbm), ptr(perm))
    return pbm, perm, outs


def candidates(meta, pbm, mode, x, y, z=0.0):
    pbm = _u32(pbm)
    c = lib().orc_spatial_candidates(ctypes.byref(meta), ptr(pbm), mode, x, y, z, None, 0)
    out = np.zeros(max(c, 1), dtype=np.uint32)
    lib().orc_spatial_candidates(ctypes.byref(meta), ptr(pbm), mode, x, y, z, ptr(out), c)
    return out[:c]


def mean_candidates(meta, pbm, mode=0):
    pbm = _u32(pbm)
    return lib().orc_spatial_mean_candidates(ctypes.byref(meta), ptr(pbm), mode)


def compact(flags, keep, arrays):
    flags = _u32(flags)
    n = len(flags)
    outs = [np.zeros_like(a) for a in arrays]
    ins = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    outp = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in outs])
    nb = (ctypes.c_uint32 * len(arrays))(*[a.itemsize for a in arrays])
    c = lib().orc_compact(ctypes.c_uint32(n), ptr(flags), ctypes.c_uint32(keep), ctypes.c_uint32(len(arrays)),
                          ins, nb, outp)
    return c, [o[:c] for o in outs]


def births(flags, newbuf, dst, dst_count, id_var, next_id, by_parent_index=False):
    """dst: list of arrays with spare capacity; newbuf[id_var] is ignored."""
    flags = _u32(flags)
    nv = len(dst)
    nbp = (ctypes.c_void_p * nv)(*[(a.ctypes.data if a is not None else 0) for a in newbuf])
    dp = (ctypes.c_void_p * nv)(*[a.ctypes.data for a in dst])
    by = (ctypes.c_uint32 * nv)(*[a.itemsize for a in dst])
    nid = ctypes.c_uint32(next_id)
    c = lib().orc_births_ex(ctypes.c_uint32(len(flags)), ptr(flags), ctypes.c_uint32(nv), nbp, by, dp,
                            ctypes.c_uint32(dst_count), ctypes.c_uint32(id_var), ctypes.byref(nid),
                            ctypes.c_int32(1 if by_parent_index else 0))
    return c, nid.value


def agent_sort_keys(dims, mn, mx, radius, x, y, z=None):
    x = _f32(x); y = _f32(y)
    n = len(x)
    keys = np.zeros(n, dtype=np.uint32)
    bits = ctypes.c_uint32(0)
    mnp = (ctypes.c_float * 3)(*(list(mn) + [0.0] * (3 - len(mn))))
    mxp = (ctypes.c_float * 3)(*(list(mx) + [0.0] * (3 - len(mx))))
    zarr = _f32(z) if z is not None else None
    lib().orc_agent_sort_keys(dims, mnp, mxp, ctypes.c_float(radius), ctypes.c_uint32(n), ptr(x, ctypes.c_float),
                              ptr(y, ctypes.c_float), ptr(zarr, ctypes.c_float) if z is not None else None,
                              ptr(keys), ctypes.byref(bits))
    return keys, bits.value


def stable_sort_perm(keys):
    keys = _u32(keys)
    perm = np.zeros(len(keys), dtype=np.uint32)
    lib().orc_stable_sort_perm(ctypes.c_uint32(len(keys)), ptr(keys), ptr(perm))
    return perm


class Sim:
    """Whole-model oracle simulation (model ids: 0 circles, 1 boids, 2 predator-prey, 3 circles brute force)."""

    def __init__(self, model, params, seed, bin_mult=1.0):
        p = _f32(params)
        self.h = lib().orc_sim_create(model, ptr(p, ctypes.c_float), len(p), seed)
        assert self.h, "bad oracle model"
        if bin_mult != 1.0:
            assert lib().orc_sim_set_bin_mult(self.h, bin_mult) == 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def set_agents(self, t, arrays, ids=None):
        n = len(arrays[0])
        arrays = [np.ascontiguousarray(a) for a in arrays]
        assert len(arrays) == lib().orc_sim_var_count(self.h, t)
        vp = (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
        idp = ptr(_u32(ids)) if ids is not None else None
        if ids is not None:
            ida = _u32(ids)
            idp = ptr(ida)
        assert lib().orc_sim_set_agents(self.h, t, n, vp, idp) == 0

    def count(self, t=0):
        return lib().orc_sim_count(self.h, t)

    def get_agents(self, t, dtypes):
        n = self.count(t)
        outs = [np.zeros(n, dtype=d) for d in dtypes]
        ids = np.zeros(n, dtype=np.uint32)
        vp = (ctypes.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
        assert lib().orc_sim_get_agents(self.h, t, vp, ptr(ids)) == 0
        return outs, ids

    def step(self, n=1):
        assert lib().orc_sim_step(self.h, n) == 0

    def pbm(self, msg=0):
        b = lib().orc_sim_bins(self.h, msg)
        out = np.zeros(b + 1, dtype=np.uint32)
        assert lib().orc_sim_get_pbm(self.h, msg, ptr(out)) == 0
        return out

    def last_candidates(self, msg=0):
        return lib().orc_sim_last_candidates(self.h, msg)
