"""ctypes binding of libfgx.so (include/fgx.h).

Device buffers are passed as integer device addresses (torch CUDA tensors via
.data_ptr()); streams are hipStream_t handles as integers (0 = null stream).
Every non-zero status raises FgxError carrying fgx_last_error().
"""
import ctypes
import os

from . import load

_c = ctypes
vp = _c.c_void_p
u32 = _c.c_uint32
i32 = _c.c_int32
sz = _c.c_size_t

FGX_OP_SPATIAL_BUILD = 0
FGX_OP_COMPACT = 1
FGX_OP_BIRTHS = 2
FGX_OP_AGENT_SORT = 3
FGX_OP_SCAN = 4
FGX_OP_STATS = 5
FGX_MAX_VARS = 32
FGX_OPT_KEEP_SCRATCH_ZERO = 1


class FgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fgx error {code}: {msg}")
        self.code = code


class SpatialMeta(_c.Structure):
    _fields_ = [("dims", i32), ("min", _c.c_float * 3), ("max", _c.c_float * 3), ("radius", _c.c_float),
                ("width", _c.c_float * 3), ("grid", u32 * 3), ("bins", u32), ("wrap_ok", i32), ("cell", _c.c_float),
                ("span", i32)]


class SegKey(_c.Structure):
    _fields_ = [("elem_off", u32), ("n", u32), ("key_off", u32), ("keys", u32), ("tile_off", u32), ("dims", i32),
                ("min", _c.c_float * 3), ("radius", _c.c_float), ("width", _c.c_float * 3), ("grid", u32 * 3),
                ("key_mask", u32), ("pad", u32 * 3)]


FGX_SEG_MESSAGE_BINS = 0
FGX_SEG_AGENT_SORT = 1

_SIGS = {
    "fgx_ctx_create": [_c.c_int, _c.POINTER(vp)],
    "fgx_ctx_destroy": [vp],
    "fgx_device_info": [vp, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int), _c.POINTER(sz)],
    "fgx_malloc": [vp, sz, _c.POINTER(vp)],
    "fgx_free": [vp, vp],
    "fgx_memcpy_h2d": [vp, vp, sz, vp],
    "fgx_memcpy_d2h": [vp, vp, sz, vp],
    "fgx_memcpy_d2d": [vp, vp, sz, vp],
    "fgx_fill_u32": [vp, u32, sz, vp],
    "fgx_stream_synchronize": [vp],
    "fgx_stream_create": [_c.POINTER(vp)],
    "fgx_stream_destroy": [vp],
    "fgx_graph_begin": [vp],
    "fgx_graph_end": [vp, _c.POINTER(vp)],
    "fgx_graph_launch": [vp, vp],
    "fgx_graph_destroy": [vp],
    "fgx_device_synchronize": [],
    "fgx_scratch_bytes": [_c.c_int, u32, u32, _c.POINTER(sz)],
    "fgx_scratch_clear": [vp, sz, vp],
    "fgx_spatial_meta_init": [_c.c_int, _c.POINTER(_c.c_float), _c.POINTER(_c.c_float), _c.c_float,
                              _c.POINTER(SpatialMeta)],
    "fgx_ctx_set_option": [vp, _c.c_int, _c.c_int],
    "fgx_spatial_meta_init_mult": [_c.c_int, _c.POINTER(_c.c_float), _c.POINTER(_c.c_float), _c.c_float, _c.c_float,
                                   _c.POINTER(SpatialMeta)],
    "fgx_spatial_build": [vp, _c.POINTER(SpatialMeta), vp, u32, vp, vp, vp, u32, _c.POINTER(vp), _c.POINTER(vp),
                          _c.POINTER(u32), vp, vp, sz, vp],
    "fgx_spatial_build_noalloc": [vp, _c.POINTER(SpatialMeta), vp, u32, vp, vp, vp, u32, _c.POINTER(vp),
                                  _c.POINTER(vp), _c.POINTER(u32), vp, vp],
    "fgx_bucket_build": [vp, i32, u32, vp, u32, vp, u32, _c.POINTER(vp), _c.POINTER(vp), _c.POINTER(u32), vp, vp, sz,
                         vp],
    "fgx_agent_sort_key_bits": [_c.c_int, _c.POINTER(_c.c_float), _c.POINTER(_c.c_float), _c.c_float,
                                _c.POINTER(u32)],
    "fgx_agent_sort": [vp, _c.c_int, _c.POINTER(_c.c_float), _c.POINTER(_c.c_float), _c.c_float, vp, u32, vp, vp,
                       vp, u32, _c.POINTER(vp), _c.POINTER(vp), _c.POINTER(u32), vp, sz, vp],
    "fgx_compact": [vp, vp, u32, vp, u32, u32, _c.POINTER(vp), _c.POINTER(vp), _c.POINTER(u32), vp, vp, sz, vp],
    "fgx_append_births": [vp, vp, vp, u32, u32, _c.POINTER(vp), _c.POINTER(vp), _c.POINTER(u32), i32, vp, u32, vp,
                          vp, vp, vp, sz, vp],
    "fgx_exclusive_scan_u32": [vp, vp, vp, u32, vp, sz, vp],
    "fgx_assign_ids": [vp, u32, u32, vp],
    "fgx_reduce_sum": [vp, _c.c_int, vp, vp, u32, vp, vp, sz, vp],
    "fgx_reduce_stats": [vp, _c.c_int, vp, vp, u32, vp, vp, sz, vp],
    "fgx_seg_key_init": [_c.c_int, _c.c_int, _c.POINTER(_c.c_float), _c.POINTER(_c.c_float), _c.c_float,
                         _c.POINTER(SegKey)],
    "fgx_segments_pack": [_c.POINTER(SegKey), u32, _c.POINTER(u32), _c.POINTER(u32), _c.POINTER(u32)],
    "fgx_spatial_build_seg": [vp, vp, u32, u32, u32, vp, u32, vp, vp, vp, u32, _c.POINTER(vp), _c.POINTER(vp),
                              _c.POINTER(u32), vp, vp, sz, vp],
    "fgx_agent_sort_seg": [vp, vp, u32, u32, u32, vp, u32, vp, vp, vp, u32, _c.POINTER(vp), _c.POINTER(vp),
                           _c.POINTER(u32), vp, sz, vp],
    "fgx_seg_layout": [u32, vp, vp, vp, vp, vp, u32, vp],
    "fgx_seg_count": [u32, vp, vp, vp, u32, vp, vp],
    "fgx_seg_append": [u32, vp, vp, vp, u32, u32, _c.POINTER(vp), _c.POINTER(vp), _c.POINTER(vp), _c.POINTER(u32),
                       _c.c_int32, vp, vp, vp],
    "fgx_add_u32_n": [vp, u32, u32, vp],
    "fgx_pack_records": [vp, u32, vp, vp, vp, vp, vp, vp],
    "fgx_spatial_build_records": [vp, _c.POINTER(SpatialMeta), vp, u32, vp, vp, vp, u32, _c.POINTER(vp),
                                  _c.POINTER(vp), _c.POINTER(u32), _c.POINTER(_c.c_int32), vp, vp, vp, sz, vp],
    "fgx_copy_ranges": [u32, _c.POINTER(vp), _c.POINTER(vp), _c.POINTER(u32), vp, u32, u32, vp],
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = load("libfgx.so")
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _c.c_int
        L.fgx_last_error.restype = _c.c_char_p
        L.fgx_last_error.argtypes = []
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise FgxError(rc, lib().fgx_last_error().decode())
    return rc


def header_symbols():
    """Every function declared in include/fgx.h."""
    import re
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    txt = open(os.path.join(here, "include", "fgx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fgx_\w+)\s*\(", txt, flags=re.M)))


class Context:
    def __init__(self, device=0):
        h = vp()
        check(lib().fgx_ctx_create(device, _c.byref(h)))
        self.h = h

    def set_option(self, option, value):
        check(lib().fgx_ctx_set_option(self.h, option, int(value)))

    def close(self):
        if self.h:
            lib().fgx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _arr(t, vals):
    return (t * max(len(vals), 1))(*vals)


def spatial_meta(dims, mn, mx, radius, mult=1.0):
    m = SpatialMeta()
    mn = list(mn) + [0.0] * (3 - len(mn))
    mx = list(mx) + [0.0] * (3 - len(mx))
    check(lib().fgx_spatial_meta_init_mult(dims, _arr(_c.c_float, mn), _arr(_c.c_float, mx), radius, mult,
                                           _c.byref(m)))
    return m


def scratch_bytes(op, n, bins=0):
    b = sz()
    check(lib().fgx_scratch_bytes(op, n, bins, _c.byref(b)))
    return b.value


def spatial_build(ctx, meta, n_cap, x, y, z, vars_in, vars_out, var_bytes, pbm, scratch, scratch_len,
                  stream=0, d_n=None):
    nv = len(vars_in)
    check(lib().fgx_spatial_build(ctx.h, _c.byref(meta), d_n, n_cap, x, y, z, nv, _arr(vp, vars_in),
                                  _arr(vp, vars_out), _arr(u32, var_bytes), pbm, scratch, scratch_len, stream))


def spatial_build_records(ctx, meta, n_cap, x, y, z, vars_in, vars_out, var_bytes, rec_vars, records, pbm, scratch,
                          scratch_len, stream=0, d_n=None):
    """fgx_spatial_build that also writes 16-byte records of vars rec_vars[0..3] (-1: 0 word) in bin order."""
    nv = len(vars_in)
    check(lib().fgx_spatial_build_records(ctx.h, _c.byref(meta), d_n, n_cap, x, y, z, nv, _arr(vp, vars_in),
                                          _arr(vp, vars_out), _arr(u32, var_bytes), _arr(_c.c_int32, rec_vars),
                                          records, pbm, scratch, scratch_len, stream))


def spatial_build_noalloc(ctx, meta, n_cap, x, y, z, vars_in, vars_out, var_bytes, pbm, stream=0, d_n=None):
    """fgx_spatial_build working in the library's own per-GPU space (no scratch argument)."""
    nv = len(vars_in)
    check(lib().fgx_spatial_build_noalloc(ctx.h, _c.byref(meta), d_n, n_cap, x, y, z, nv, _arr(vp, vars_in),
                                          _arr(vp, vars_out), _arr(u32, var_bytes), pbm, stream))


def bucket_build(ctx, key_min, keys, n_cap, key, vars_in, vars_out, var_bytes, pbm, scratch, scratch_len, stream=0,
                 d_n=None):
    nv = len(vars_in)
    check(lib().fgx_bucket_build(ctx.h, key_min, keys, d_n, n_cap, key, nv, _arr(vp, vars_in), _arr(vp, vars_out),
                                 _arr(u32, var_bytes), pbm, scratch, scratch_len, stream))


def agent_sort(ctx, dims, mn, mx, radius, n_cap, x, y, z, vars_in, vars_out, var_bytes, scratch, scratch_len,
               stream=0, d_n=None):
    mn = list(mn) + [0.0] * (3 - len(mn))
    mx = list(mx) + [0.0] * (3 - len(mx))
    nv = len(vars_in)
    check(lib().fgx_agent_sort(ctx.h, dims, _arr(_c.c_float, mn), _arr(_c.c_float, mx), radius, d_n, n_cap, x, y,
                               z, nv, _arr(vp, vars_in), _arr(vp, vars_out), _arr(u32, var_bytes), scratch,
                               scratch_len, stream))


def agent_sort_key_bits(dims, mn, mx, radius):
    mn = list(mn) + [0.0] * (3 - len(mn))
    mx = list(mx) + [0.0] * (3 - len(mx))
    b = u32()
    check(lib().fgx_agent_sort_key_bits(dims, _arr(_c.c_float, mn), _arr(_c.c_float, mx), radius, _c.byref(b)))
    return b.value


def compact(ctx, flags, keep, n_cap, vars_in, vars_out, var_bytes, d_count_out, scratch, scratch_len, stream=0,
            d_n=None):
    nv = len(vars_in)
    check(lib().fgx_compact(ctx.h, flags, keep, d_n, n_cap, nv, _arr(vp, vars_in), _arr(vp, vars_out),
                            _arr(u32, var_bytes), d_count_out, scratch, scratch_len, stream))


def append_births(ctx, flags, n_parents_cap, newbuf, dst, var_bytes, id_var, d_dst_count, dst_capacity, d_next_id,
                  scratch, scratch_len, stream=0, d_n_parents=None, d_overflow=None, d_id_slots=None):
    nv = len(dst)
    check(lib().fgx_append_births(ctx.h, flags, d_n_parents, n_parents_cap, nv, _arr(vp, newbuf), _arr(vp, dst),
                                  _arr(u32, var_bytes), id_var, d_dst_count, dst_capacity, d_next_id, d_id_slots,
                                  d_overflow, scratch, scratch_len, stream))


def exclusive_scan(ctx, d_in, d_out, n, scratch, scratch_len, stream=0):
    check(lib().fgx_exclusive_scan_u32(ctx.h, d_in, d_out, n, scratch, scratch_len, stream))


def fill_u32(dst, value, n, stream=0):
    check(lib().fgx_fill_u32(dst, value, n, stream))


def scratch_clear(scratch, nbytes, stream=0):
    check(lib().fgx_scratch_clear(scratch, nbytes, stream))


def reduce_sum(ctx, type_code, d_in, n_cap, d_out, scratch, scratch_len, stream=0, d_n=None):
    check(lib().fgx_reduce_sum(ctx.h, type_code, d_in, d_n, n_cap, d_out, scratch, scratch_len, stream))


def reduce_stats(ctx, type_code, d_in, n_cap, d_out, scratch, scratch_len, stream=0, d_n=None):
    """Four 8-byte words at d_out: sum, min, max (double or 64-bit integer) and the sum of
    squared deviations from the mean (double)."""
    check(lib().fgx_reduce_stats(ctx.h, type_code, d_in, d_n, n_cap, d_out, scratch, scratch_len, stream))


def seg_key(kind, dims, mn, mx, radius, n=0):
    k = SegKey()
    mn = list(mn) + [0.0] * (3 - len(mn))
    mx = list(mx) + [0.0] * (3 - len(mx))
    check(lib().fgx_seg_key_init(kind, dims, _arr(_c.c_float, mn), _arr(_c.c_float, mx), radius, _c.byref(k)))
    k.n = n
    return k


def segments_pack(keys):
    """Fills the offsets of a list of SegKey in place; returns (n_total, keys_total, tiles_total, ctypes array)."""
    arr = (SegKey * max(len(keys), 1))(*keys)
    n, kt, t = u32(), u32(), u32()
    check(lib().fgx_segments_pack(arr, len(keys), _c.byref(n), _c.byref(kt), _c.byref(t)))
    return n.value, kt.value, t.value, arr


def spatial_build_seg(ctx, d_segs, nsegs, tiles, n_total, keys_total, x, y, z, vars_in, vars_out, var_bytes, pbm,
                      scratch, scratch_len, stream=0, d_n_total=None):
    nv = len(vars_in)
    check(lib().fgx_spatial_build_seg(ctx.h, d_segs, nsegs, tiles, n_total, d_n_total, keys_total, x, y, z, nv,
                                      _arr(vp, vars_in), _arr(vp, vars_out), _arr(u32, var_bytes), pbm, scratch,
                                      scratch_len, stream))


def agent_sort_seg(ctx, d_segs, nsegs, tiles, n_total, keys_total, x, y, z, vars_in, vars_out, var_bytes, scratch,
                   scratch_len, stream=0, d_n_total=None):
    nv = len(vars_in)
    check(lib().fgx_agent_sort_seg(ctx.h, d_segs, nsegs, tiles, n_total, d_n_total, keys_total, x, y, z, nv,
                                   _arr(vp, vars_in), _arr(vp, vars_out), _arr(u32, var_bytes), scratch, scratch_len,
                                   stream))


class SegTarget(_c.Structure):
    _fields_ = [("table", vp), ("stride", u32), ("off_field", _c.c_int32), ("n_field", _c.c_int32),
                ("tile_field", _c.c_int32), ("tile_elems", u32), ("sentinel", _c.c_int32)]


def seg_layout(nsegs, d_cnt, d_off, d_total, targets=(), d_add=None, stream=0):
    arr = (SegTarget * max(len(targets), 1))(*targets)
    check(lib().fgx_seg_layout(nsegs, d_cnt, d_add, d_off, d_total, arr, len(targets), stream))


def seg_count(nsegs, d_off, d_cnt, flags, keep, d_out, stream=0):
    check(lib().fgx_seg_count(nsegs, d_off, d_cnt, flags, keep, d_out, stream))


def seg_append(nsegs, d_off, d_cnt, d_born, n_cap_out, list_in, born_in, list_out, var_bytes, id_var, d_next_id, d_work,
               stream=0):
    check(lib().fgx_seg_append(nsegs, d_off, d_cnt, d_born, n_cap_out, len(list_in), _arr(vp, list_in),
                               _arr(vp, born_in), _arr(vp, list_out), _arr(u32, var_bytes), id_var, d_next_id, d_work,
                               stream))


def add_u32_n(p, n, delta, stream=0):
    check(lib().fgx_add_u32_n(p, n, delta, stream))


def stream_create():
    h = vp()
    check(lib().fgx_stream_create(_c.byref(h)))
    return h.value


def stream_destroy(stream):
    check(lib().fgx_stream_destroy(stream))


def stream_synchronize(stream):
    check(lib().fgx_stream_synchronize(stream))


def graph_capture(stream, enqueue):
    """Captures what enqueue() puts on `stream` into a graph; returns the graph handle."""
    check(lib().fgx_graph_begin(stream))
    try:
        enqueue()
    finally:
        g = vp()
        rc = lib().fgx_graph_end(stream, _c.byref(g))
    check(rc)
    return g.value


def graph_launch(graph, stream):
    check(lib().fgx_graph_launch(graph, stream))


def graph_destroy(graph):
    check(lib().fgx_graph_destroy(graph))


def pack_records(d_n, n_cap, w0, w1, w2, w3, out, stream=0):
    """out[i] = {w0[i], w1[i], w2[i], w3[i]} (16-byte records; None/0 sources give 0 words)."""
    check(lib().fgx_pack_records(d_n, n_cap, w0, w1, w2, w3, out, stream))


def copy_ranges(src, dst, var_bytes, d_ranges, nranges, n_total, stream=0):
    check(lib().fgx_copy_ranges(len(src), _arr(vp, src), _arr(vp, dst), _arr(u32, var_bytes), d_ranges, nranges, n_total,
                                stream))
