"""Pins the serial CPU oracle before anything is compared against it.

Each check is independent of the oracle's own formulas: published known-answer
vectors, a second implementation (rocRAND), textbook closed forms computed
with numpy, brute-force enumeration, and restatements of the reference's own
known-answer tests (tests/test_cases/runtime/messaging/test_spatial_3d.cu etc.).
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- RNG
def test_philox_published_kat():
    d = json.load(open(os.path.join(GOLD, "philox_random123_kat.json")))
    for v in d["vectors"]:
        ctr = [int(h, 16) for h in v["ctr"]]
        key = [int(h, 16) for h in v["key"]]
        assert O.philox(ctr, key) == [int(h, 16) for h in v["out"]]


def test_philox_matches_rocrand():
    n = 0
    for line in open(os.path.join(GOLD, "philox_rocrand.jsonl")):
        d = json.loads(line)
        assert O.philox(d["ctr"], d["key"]) == d["out"], d
        n += 1
    assert n >= 60


def test_agent_stream_layout_and_uniform():
    L = O.lib()
    seed = 0x1234_5678_9ABC_DEF0
    for (aid, step, stream) in [(1, 0, 0), (77, 5, 3), (2**32 - 1, 999, 8)]:
        for k in range(9):
            blk = O.philox([k >> 2, stream, step, aid], [seed & 0xFFFFFFFF, seed >> 32])
            u = L.orc_rng_u32(seed, aid, step, stream, k)
            assert u == blk[k & 3]
            f = L.orc_rng_uniform(seed, aid, step, stream, k)
            # AgentRandom.cuh:80-83: low 24 bits scaled by 2^-24, range [0,1)
            assert f == np.float32((u & 0xFFFFFF) / 16777216.0)
            assert 0.0 <= f < 1.0


# ---------------------------------------------------------------- grid metadata
@pytest.mark.parametrize("dims,width,radius,grid,wrap", [
    (3, np.float32(np.cbrt(16384.0)), 2.0, 13, False),   # circles: last bin 1.398 wide
    (3, 3.2, 0.05, 64, True),                            # boids: 3.2f/0.05f == 64 exactly
    (2, 1000.0, 5.0, 200, True),                         # predator-prey
    (3, 50.1, 10.0, 6, False),                           # test_spatial_3d.cu:1111 bounds_not_factor_radius
    (3, 5.0, 1.0, 5, True),                              # test_spatial_3d.cu:72 Mandatory
    (3, 70.0, 3.5, 20, True),                            # test_spatial_3d.cu:898 Wrapped
])
def test_grid_dims(dims, width, radius, grid, wrap):
    m = O.spatial_meta(dims, [0.0] * dims, [float(width)] * dims, radius)
    assert list(m.grid)[:dims] == [grid] * dims
    assert m.bins == grid ** dims
    assert bool(m.wrap_ok) == wrap


def _np_keys(meta, x, y, z=None):
    """Bin index closed form (float32 arithmetic, clamped, x fastest)."""
    r = np.float32(meta.radius)
    def cell(p, a):
        g = np.floor((np.float32(p) - np.float32(meta.min[a])) / r).astype(np.int64)
        return np.clip(g, 0, meta.grid[a] - 1)
    k = cell(x, 0) + cell(y, 1) * meta.grid[0]
    if z is not None:
        k = k + cell(z, 2) * meta.grid[0] * meta.grid[1]
    return k.astype(np.int64)


@pytest.mark.parametrize("dims,n,width,radius,seed", [
    (3, 16384, np.float32(np.cbrt(16384.0)), 2.0, 12),
    (3, 20000, 3.2, 0.05, 7),
    (2, 30000, 1000.0, 5.0, 3),
    (3, 1, 5.0, 1.0, 1),
    (3, 0, 5.0, 1.0, 1),
])
def test_spatial_build_matches_counting_sort(dims, n, width, radius, seed):
    rng = np.random.default_rng(seed)
    P = rng.uniform(-0.2 * width, 1.2 * width, size=(dims, n)).astype(np.float32)  # includes out-of-range
    m = O.spatial_meta(dims, [0.0] * dims, [float(width)] * dims, radius)
    ids = np.arange(n, dtype=np.uint32) + 1
    pbm, perm, outs = O.spatial_build(m, P[0], P[1], P[2] if dims == 3 else None, extra=[ids])
    keys = _np_keys(m, P[0], P[1], P[2] if dims == 3 else None)
    hist = np.bincount(keys, minlength=m.bins)
    expect_pbm = np.concatenate([[0], np.cumsum(hist)]).astype(np.uint32)
    assert np.array_equal(pbm, expect_pbm)
    assert pbm[-1] == n and np.all(np.diff(pbm.astype(np.int64)) >= 0)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, order)
    assert np.array_equal(outs[-1], ids[order])
    assert np.array_equal(outs[0], P[0][order])


def _bin_coords(meta, p):
    r = np.float32(meta.cell)
    return [np.clip(np.floor((np.float32(p[a]) - np.float32(meta.min[a])) / r).astype(np.int64), 0, meta.grid[a] - 1)
            for a in range(meta.dims)]


def _check_filter_against_bruteforce(dims, n, width, radius, seed, nq=300, mult=1.0):
    rng = np.random.default_rng(seed)
    P = rng.uniform(0, width, size=(dims, n)).astype(np.float32)
    m = O.spatial_meta(dims, [0.0] * dims, [float(width)] * dims, radius, mult)
    S = m.span
    pbm, perm, _ = O.spatial_build(m, P[0], P[1], P[2] if dims == 3 else None)
    sortedP = P[:, perm]
    cells = np.stack(_bin_coords(m, sortedP))
    for q in rng.choice(n, size=min(nq, n), replace=False):
        qp = P[:, q]
        qc = np.array([c[0] for c in _bin_coords(m, qp[:, None])])
        got = O.candidates(m, pbm, 0, *[float(v) for v in qp])
        # brute force: every message whose (clamped) bin is within +-S on every axis
        mask = np.all(np.abs(cells - qc[:, None]) <= S, axis=0)
        assert sorted(got.tolist()) == np.nonzero(mask)[0].tolist()
        # visit order: strips (dy outer, dz inner), ascending x then stable within bin
        if dims == 3:
            strip_key = (cells[1, got] - qc[1]) * (2 * S + 1) + (cells[2, got] - qc[2])
        else:
            strip_key = cells[1, got] - qc[1]
        assert np.all(np.diff(strip_key) >= 0)
        assert np.all(np.diff(got.astype(np.int64))[np.diff(strip_key) == 0] > 0)
        # superset of the in-radius set
        d = np.sqrt(((sortedP - qp[:, None]).astype(np.float64) ** 2).sum(axis=0))
        assert set(np.nonzero(d < radius)[0].tolist()) <= set(got.tolist())


def test_filter_3d_mandatory_known_answer():
    # test_spatial_3d.cu:72-205: 2049 agents in [0,5)^3, r=1, count = sum of clamped 27-neighbourhood
    _check_filter_against_bruteforce(3, 2049, 5.0, 1.0, seed=1)


def test_filter_2d_mandatory_known_answer():
    # test_spatial_2d.cu:66 (2-D twin)
    _check_filter_against_bruteforce(2, 2049, 5.0, 1.0, seed=2)


def test_filter_not_factor_grid():
    _check_filter_against_bruteforce(3, 4000, float(np.float32(np.cbrt(16384.0))), 2.0, seed=3)


def test_bounds_not_factor_radius_known_answer():
    # test_spatial_3d.cu:1111-1189: IDs 1-4 see 2 messages, IDs 5-6 see 1
    h = np.float32(50.1) / np.float32(2)
    pts = np.array([[h, 0.0, h], [h, 18.0, h], [0.0, h, 0.5], [18.0, h, 0.5],
                    [h, 50.0, 50.1], [h, 50.0, np.float32(50.1) - np.float32(10.11)]], dtype=np.float32).T
    m = O.spatial_meta(3, [0, 0, 0], [50.1, 50.1, 50.1], 10.0)
    pbm, perm, _ = O.spatial_build(m, pts[0], pts[1], pts[2])
    counts = [len(O.candidates(m, pbm, 0, *[float(v) for v in pts[:, i]])) for i in range(6)]
    assert counts == [2, 2, 2, 2, 1, 1]


@pytest.mark.parametrize("offset", [(0.0, 0.0, 0.0), (141.0, -540.0, 200.0), (-1401.5, 5640.3, -2008.8)])
def test_wrapped_3d_known_answer(offset):
    # test_spatial_3d.cu:855-975: 35^3 lattice spacing 2 in a 70-wide env, r=3.5: within the
    # radius every agent sees exactly 27 messages (incl. itself), virtual displacements sum to 0.
    off = np.array(offset, dtype=np.float32)
    i, j, k = np.meshgrid(np.arange(35), np.arange(35), np.arange(35), indexing="ij")
    P = np.stack([(i.ravel() * np.float32(2.0)).astype(np.float32) + off[0],
                  (j.ravel() * np.float32(2.0)).astype(np.float32) + off[1],
                  (k.ravel() * np.float32(2.0)).astype(np.float32) + off[2]]).astype(np.float32)
    mn = [float(o) for o in off]
    mx = [float(np.float32(70.0) + o) for o in off]
    m = O.spatial_meta(3, mn, mx, 3.5)
    assert m.wrap_ok
    pbm, perm, _ = O.spatial_build(m, P[0], P[1], P[2])
    S = P[:, perm]
    width = np.array(list(m.width), dtype=np.float32)
    rng = np.random.default_rng(0)
    sample = np.concatenate([[0, len(perm) - 1], rng.choice(len(perm), 400, replace=False)])
    for q in sample:
        qp = S[:, q]
        got = O.candidates(m, pbm, 1, *[float(v) for v in qp])
        x2 = S[:, got]
        x21 = (x2 - qp[:, None]).astype(np.float32)
        # getVirtualX (MessageSpatial3DDevice.cuh:373-397)
        v = np.where(np.abs(x21) > width[:, None] / np.float32(2.0),
                     x2 - (x21 / np.abs(x21) * width[:, None]), x2).astype(np.float32)
        d21 = (v - qp[:, None]).astype(np.float32)
        dist = np.sqrt((d21 * d21).sum(axis=0, dtype=np.float32))
        ok = (dist <= np.float32(3.5)) & np.all((np.abs(d21) == 2.0) | (v == qp[:, None]), axis=0)
        assert ok.sum() == 27
        assert np.all(d21[:, ok].sum(axis=1) == 0)


@pytest.mark.parametrize("dims", [2, 3])
def test_filter_half_width_bins(dims):
    """Bin-width multiplier 1/2 (no reference counterpart): bins of width r/2, searched 2
    bins either side, same strip order; still a superset of the in-radius set."""
    m = O.spatial_meta(dims, [0.0] * dims, [5.0] * dims, 1.0, 0.5)
    assert m.span == 2 and m.cell == 0.5 and list(m.grid)[:dims] == [10] * dims
    _check_filter_against_bruteforce(dims, 3000, 5.0, 1.0, seed=4, nq=200, mult=0.5)
    _check_filter_against_bruteforce(dims, 3000, float(np.float32(np.cbrt(16384.0))), 2.0, seed=5, nq=100, mult=0.5)


def test_half_width_grid_dims():
    # ceil(width / (r/2)) in FP32: 25.398/1 -> 26 (= 2 x 13); 12.3/0.5 -> 25 (not 2 x ceil(12.3) = 26)
    m = O.spatial_meta(3, [0, 0, 0], [float(np.float32(np.cbrt(16384.0)))] * 3, 2.0, 0.5)
    assert list(m.grid) == [26, 26, 26]
    m = O.spatial_meta(3, [0, 0, 0], [12.3, 12.3, 12.3], 1.0, 0.5)
    assert list(m.grid) == [25, 25, 25] and not m.wrap_ok


def test_wrap_half_width_order():
    m = O.spatial_meta(3, [0, 0, 0], [4, 4, 4], 1.0, 0.5)
    assert m.wrap_ok and list(m.grid) == [8, 8, 8]
    c = np.arange(8, dtype=np.float32) * np.float32(0.5) + np.float32(0.25)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    pbm, perm, _ = O.spatial_build(m, X.ravel(), Y.ravel(), Z.ravel())
    got = O.candidates(m, pbm, 1, 0.25, 0.25, 0.25)
    r = range(-2, 3)
    expect = [((dz % 8) * 64 + (dy % 8) * 8 + (dx % 8)) for dx in r for dy in r for dz in r]
    assert got.tolist() == expect


def test_wrap_order_is_dx_dy_dz():
    m = O.spatial_meta(3, [0, 0, 0], [4, 4, 4], 1.0)
    # one message at each bin centre: its sorted index equals its bin hash
    c = np.arange(4, dtype=np.float32) + np.float32(0.5)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    pbm, perm, _ = O.spatial_build(m, X.ravel(), Y.ravel(), Z.ravel())
    got = O.candidates(m, pbm, 1, 0.5, 0.5, 0.5)
    expect = [((dz % 4) * 16 + (dy % 4) * 4 + (dx % 4)) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
    assert got.tolist() == expect


# ---------------------------------------------------------------- compaction / births / sort
def test_compact_is_stable_partition():
    rng = np.random.default_rng(5)
    n = 10007
    flags = (rng.uniform(size=n) < 0.7).astype(np.uint32)
    a = rng.uniform(size=n).astype(np.float32)
    b = np.arange(n, dtype=np.uint32)
    c, (oa, ob) = O.compact(flags, 1, [a, b])
    assert c == flags.sum()
    assert np.array_equal(oa, a[flags == 1]) and np.array_equal(ob, b[flags == 1])
    c0, (_, ob0) = O.compact(flags, 0, [a, b])
    assert c0 == n - c and np.array_equal(ob0, b[flags == 0])
    assert O.compact(np.zeros(0, np.uint32), 1, [np.zeros(0, np.float32)])[0] == 0


def test_births_append_with_scan_ids():
    rng = np.random.default_rng(6)
    n_par, existing = 5000, 1234
    flags = (rng.uniform(size=n_par) < 0.3).astype(np.uint32)
    newx = rng.uniform(size=n_par).astype(np.float32)
    cap = existing + n_par
    dstx = np.zeros(cap, np.float32); dstx[:existing] = 7.0
    dstid = np.zeros(cap, np.uint32); dstid[:existing] = np.arange(existing) + 1
    c, nid = O.births(flags, [newx, None], [dstx, dstid], existing, 1, existing + 1)
    nb = int(flags.sum())
    assert c == existing + nb and nid == existing + 1 + nb
    assert np.array_equal(dstx[existing:c], newx[flags == 1])
    assert np.array_equal(dstid[existing:c], np.arange(existing + 1, existing + 1 + nb))
    assert np.all(dstx[:existing] == 7.0)
    assert len(set(dstid[:c].tolist())) == c


def test_agent_sort_reverse_known_answer():
    # test_spatial_agent_sort.cu:70-115: 4 agents at (-i,-i,-i), env [-5,5]^3, r=0.2 -> order 3,2,1,0
    x = -np.arange(4, dtype=np.float32)
    keys, bits = O.agent_sort_keys(3, [-5, -5, -5], [5, 5, 5], 0.2, x, x, x)
    assert bits == int(np.floor(np.log2(50 ** 3))) + 1
    assert O.stable_sort_perm(keys).tolist() == [3, 2, 1, 0]


def test_agent_sort_key_closed_form():
    rng = np.random.default_rng(8)
    n = 5000
    P = rng.uniform(-1, 4.5, size=(3, n)).astype(np.float32)
    keys, bits = O.agent_sort_keys(3, [0, 0, 0], [3.2, 3.2, 3.2], 0.05, P[0], P[1], P[2])
    w = np.float32(3.2); g = 64
    c = [np.floor(((P[a] - np.float32(0)) / w) * np.float32(g)).astype(np.int64) for a in range(3)]
    expect = ((c[2] * g * g + c[1] * g + c[0]) & ((1 << bits) - 1)).astype(np.uint32)
    assert np.array_equal(keys, expect)
    perm = O.stable_sort_perm(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable"))


# ---------------------------------------------------------------- whole models
def _circles(n, seed, model):
    from flamegpu2_amd import workloads
    params, arrs = workloads.circles(n, seed)
    s = O.Sim(model, params, seed)
    s.set_agents(0, arrs)
    return s, params


def test_circles_spatial_equals_bruteforce():
    # The spatial and brute-force message types must find the same in-radius neighbours;
    # only the summation order differs (a different iteration order), so FP32 agrees to ~1e-6.
    sp, _ = _circles(3000, 12, 0)
    bf, _ = _circles(3000, 12, 3)
    sp.step(1); bf.step(1)
    (a, ida) = sp.get_agents(0, [np.float32] * 4)
    (b, idb) = bf.get_agents(0, [np.float32] * 4)
    oa, ob = np.argsort(ida), np.argsort(idb)
    assert np.array_equal(ida[oa], idb[ob])
    for k in range(4):
        np.testing.assert_allclose(a[k][oa], b[k][ob], rtol=1e-5, atol=1e-6)


def test_circles_half_width_bins_match_full_bins():
    # The same in-radius neighbours are found with half-width bins; only the visit order
    # (so the FP32 summation order) differs.
    from flamegpu2_amd import workloads
    params, arrs = workloads.circles(3000, 12)
    a = O.Sim(0, params, 12)
    b = O.Sim(0, params, 12, bin_mult=0.5)
    a.set_agents(0, arrs); b.set_agents(0, arrs)
    a.step(2); b.step(2)
    (x, ia), (y, ib) = a.get_agents(0, [np.float32] * 4), b.get_agents(0, [np.float32] * 4)
    oa, ob = np.argsort(ia), np.argsort(ib)
    assert np.array_equal(ia[oa], ib[ob])
    for k in range(4):
        np.testing.assert_allclose(x[k][oa], y[k][ob], rtol=1e-5, atol=1e-6)


def test_circles_isolated_agent_stays():
    params = np.array([10.0, 2.0, 0.05], np.float32)
    s = O.Sim(0, params, 1)
    x = np.array([1.0, 8.0], np.float32)
    s.set_agents(0, [x, x.copy(), x.copy(), np.zeros(2, np.float32)])
    s.step(3)
    (a, ids) = s.get_agents(0, [np.float32] * 4)
    assert sorted(a[0].tolist()) == [1.0, 8.0]
    assert np.all(a[3] == 0)


def test_boids_isolated_agents_move_straight():
    from flamegpu2_amd import workloads
    params, _ = workloads.boids(4)
    s = O.Sim(1, params, 7)
    x = np.array([0.5, 2.0], np.float32); y = np.array([0.5, 2.5], np.float32); z = np.array([0.5, 1.0], np.float32)
    fx = np.array([0.05, 0.0], np.float32); fy = np.array([0.0, -0.05], np.float32); fz = np.zeros(2, np.float32)
    s.set_agents(0, [x, y, z, fx, fy, fz])
    s.step(4)
    (a, ids) = s.get_agents(0, [np.float32] * 6)
    o = np.argsort(ids)
    ex = x.copy(); ey = y.copy()
    for _ in range(4):
        ex = (ex + fx * np.float32(1.0)).astype(np.float32)
        ey = (ey + fy * np.float32(1.0)).astype(np.float32)
    assert np.array_equal(a[0][o], ex) and np.array_equal(a[1][o], ey)


def test_predprey_population_accounting_small():
    from flamegpu2_amd import workloads
    params, prey, pred = workloads.predprey(4000, 1000, seed=3, width=100.0)
    s = O.Sim(2, params, 3)
    s.set_agents(0, prey); s.set_agents(1, pred)
    seen_prey, seen_pred = set(range(1, 4001)), set(range(1, 1001))
    for step in range(20):
        s.step(1)
        (_, ids_prey) = s.get_agents(0, [np.float32] * 4)
        (pa, ids_pred) = s.get_agents(1, [np.float32] * 4 + [np.uint32])
        assert len(set(ids_prey.tolist())) == len(ids_prey)
        assert len(set(ids_pred.tolist())) == len(ids_pred)
        new = set(ids_prey.tolist()) - seen_prey
        # births get contiguous IDs continuing the counter
        if new:
            assert min(new) == max(seen_prey) + 1 and max(new) == max(seen_prey) + len(new)
        seen_prey |= new
        seen_pred |= set(ids_pred.tolist())
        assert np.all(pa[4] < 50)
        assert np.all((pa[0] >= 0) & (pa[0] <= 100))
