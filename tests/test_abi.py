"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/fgx.h declares, and its host-side helpers agree with the oracle."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from flamegpu2_amd import fgx, lib_path


def test_libfgx_exports_every_declared_symbol():
    syms = fgx.header_symbols()
    assert len(syms) >= 30
    L = ctypes.CDLL(lib_path("libfgx.so"))
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_model_libraries_export_model_abi():
    import glob
    libs = sorted(glob.glob(lib_path("libfgxm_*.so")))
    txt = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "fgx_model.h")).read()
    import re
    syms = sorted(set(re.findall(r"^\s*(?:int|const char\*|uint32_t|double)\s+(fgxm_\w+)\s*\(", txt, flags=re.M)))
    assert len(syms) >= 8
    assert libs, "no model libraries built"
    for p in libs:
        L = ctypes.CDLL(p)
        missing = [s for s in syms if not hasattr(L, s)]
        assert not missing, (p, missing)


@pytest.mark.parametrize("dims,mn,mx,r", [
    (3, [0, 0, 0], [25.398416, 25.398416, 25.398416], 2.0),
    (3, [0, 0, 0], [3.2, 3.2, 3.2], 0.05),
    (2, [0, 0], [1000, 1000], 5.0),
    (3, [-1401.5, 5640.3, -2008.8], [-1331.5, 5710.3, -1938.8], 3.5),
    (3, [0, 0, 0], [50.1, 50.1, 50.1], 10.0),
])
def test_spatial_meta_matches_oracle(dims, mn, mx, r):
    a = fgx.spatial_meta(dims, mn, mx, r)
    b = O.spatial_meta(dims, mn, mx, r)
    for f in ["dims", "radius", "bins", "wrap_ok"]:
        assert getattr(a, f) == getattr(b, f), f
    assert list(a.grid) == list(b.grid)
    assert list(a.width) == list(b.width)


def test_spatial_meta_rejects_bad_input():
    with pytest.raises(fgx.FgxError):
        fgx.spatial_meta(4, [0, 0, 0], [1, 1, 1], 1.0)
    with pytest.raises(fgx.FgxError):
        fgx.spatial_meta(3, [0, 0, 0], [1, 1, 1], 0.0)
    with pytest.raises(fgx.FgxError):
        fgx.spatial_meta(3, [0, 0, 0], [1, 0, 1], 1.0)


def test_agent_sort_key_bits_matches_oracle():
    x = np.zeros(1, np.float32)
    for (dims, mn, mx, r) in [(3, [-5] * 3, [5] * 3, 0.2), (3, [0] * 3, [3.2] * 3, 0.05), (2, [0, 0], [1000, 1000], 5.0)]:
        _, bits = O.agent_sort_keys(dims, mn, mx, r, x, x, x if dims == 3 else None)
        assert fgx.agent_sort_key_bits(dims, mn, mx, r) == bits


def test_scratch_sizes_monotone():
    for op in [fgx.FGX_OP_SPATIAL_BUILD, fgx.FGX_OP_COMPACT, fgx.FGX_OP_BIRTHS, fgx.FGX_OP_SCAN, fgx.FGX_OP_AGENT_SORT]:
        a = fgx.scratch_bytes(op, 1000, 100)
        b = fgx.scratch_bytes(op, 100000, 100000)
        assert 0 < a <= b and a % 256 == 0 and b % 256 == 0
    with pytest.raises(fgx.FgxError):
        fgx.scratch_bytes(99, 1, 1)


def test_seg_key_geometry_matches_single_simulation_entries():
    for dims, mn, mx, r in [(3, [0, 0, 0], [25.398416] * 3, 2.0), (2, [0, 0], [1000, 1000], 5.0),
                            (3, [0, 0, 0], [50.1] * 3, 10.0)]:
        k = fgx.seg_key(fgx.FGX_SEG_MESSAGE_BINS, dims, mn, mx, r)
        m = fgx.spatial_meta(dims, mn, mx, r)
        assert k.keys == m.bins
        assert [k.grid[a] for a in range(3)] == [max(m.grid[a], 1) for a in range(3)]
        assert [k.width[a] for a in range(dims)] == [m.width[a] for a in range(dims)]
        s = fgx.seg_key(fgx.FGX_SEG_AGENT_SORT, dims, mn, mx, r)
        bits = fgx.agent_sort_key_bits(dims, mn, mx, r)
        assert s.keys == 1 << bits and s.key_mask == (1 << bits) - 1


def test_segments_pack_offsets():
    ns = [5, 0, 2048, 2049, 1]
    keys = []
    for i, n in enumerate(ns):
        k = fgx.seg_key(fgx.FGX_SEG_MESSAGE_BINS, 3, [0, 0, 0], [4.0 + i] * 3, 1.0, n)
        keys.append(k)
    n_total, keys_total, tiles, arr = fgx.segments_pack(keys)
    assert n_total == sum(ns)
    assert keys_total == sum(k.keys for k in keys)
    assert tiles == sum((n + 2047) // 2048 for n in ns)
    assert [arr[i].elem_off for i in range(len(ns))] == list(np.cumsum([0] + ns[:-1]))
    assert [arr[i].key_off for i in range(len(ns))] == list(np.cumsum([0] + [k.keys for k in keys[:-1]]))
    assert [arr[i].tile_off for i in range(len(ns))] == list(np.cumsum([0] + [(n + 2047) // 2048 for n in ns[:-1]]))
