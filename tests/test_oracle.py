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


def test_births_ids_by_parent_index():
    # once a parent read its child's ID in-kernel, the launch reserved one ID per parent:
    # child ID = next_id + parent index, next_id advances by the parent count; children are
    # still appended in parent order
    rng = np.random.default_rng(7)
    n_par, existing = 3000, 500
    flags = (rng.uniform(size=n_par) < 0.4).astype(np.uint32)
    newx = rng.uniform(size=n_par).astype(np.float32)
    cap = existing + n_par
    dstx = np.zeros(cap, np.float32)
    dstid = np.zeros(cap, np.uint32); dstid[:existing] = np.arange(existing) + 1
    c, nid = O.births(flags, [newx, None], [dstx, dstid], existing, 1, existing + 1, by_parent_index=True)
    par = np.nonzero(flags)[0]
    assert c == existing + len(par) and nid == existing + 1 + n_par
    assert np.array_equal(dstx[existing:c], newx[par])
    assert np.array_equal(dstid[existing:c], existing + 1 + par)
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


# ---------------------------------------------------------------- model math, closed forms
# Each case is small enough that the expected state follows from the model's rules by hand
# (float64 here, FP32 in the oracle), so a dropped term, a wrong sign or a swapped operand in
# the oracle's agent functions fails one of them.  The rules: circles
# examples/cpp/circles_spatial3D/src/main.cu:13-48; boids
# examples/python_native/boids_spatial3D_wrapped/boids_spatial3D.py:123-256; predator-prey as
# specified in oracle/models.cpp.
def _agents_by_id(sim, t, dtypes):
    a, ids = sim.get_agents(t, dtypes)
    o = np.argsort(ids)
    return [x[o].astype(np.float64) if x.dtype.kind == "f" else x[o] for x in a], ids[o]


def _circle_k(d, r=2.0, repulse=0.05):
    return np.sin(d / r * 3.141 * -2.0) * repulse


@pytest.mark.parametrize("d", [0.3, 0.5, 0.9, 1.2, 1.5, 1.99])
def test_circles_two_agents_closed_form(d):
    # A at (4,5,5), B at (4+d,5,5): each moves along x by the force k(d) = sin(-2*3.141*d/r)*0.05,
    # A by +k and B by -k (repulsion for d < r/2, attraction beyond); drift = |k|.
    params = np.array([10.0, 2.0, 0.05], np.float32)
    s = O.Sim(0, params, 1)
    x = np.array([4.0, 4.0 + d], np.float32)
    c = np.array([5.0, 5.0], np.float32)
    s.set_agents(0, [x, c.copy(), c.copy(), np.zeros(2, np.float32)])
    s.step(1)
    (X, Y, Z, D), ids = _agents_by_id(s, 0, [np.float32] * 4)
    dd = float(np.float32(4.0 + d)) - 4.0
    k = _circle_k(dd)
    assert (k < 0) == (dd < 1.0)
    np.testing.assert_allclose(X, [4.0 + k, 4.0 + dd - k], rtol=0, atol=2e-6)
    np.testing.assert_allclose(D, [abs(k)] * 2, rtol=1e-5)
    assert np.all(Y == 5.0) and np.all(Z == 5.0)


def test_circles_three_agents_and_outside_radius():
    # A (5,5,5), B (5.6,5,5), C (5,5.8,5): B-C is 1.0 apart; D (8.5,5,5) is 2.9 from B, inside the
    # searched bins but outside r, so it feels nothing and moves nothing.  Forces are averaged
    # over the in-radius neighbours.
    params = np.array([10.0, 2.0, 0.05], np.float32)
    s = O.Sim(0, params, 1)
    P = np.array([[5.0, 5.0, 5.0], [5.6, 5.0, 5.0], [5.0, 5.8, 5.0], [8.5, 5.0, 5.0]], np.float32)
    s.set_agents(0, [P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy(), np.zeros(4, np.float32)])
    s.step(1)
    (X, Y, Z, D), ids = _agents_by_id(s, 0, [np.float32] * 4)
    P64 = P.astype(np.float64)
    exp = []
    for i in range(4):
        f, cnt = np.zeros(3), 0
        for j in range(4):
            v = P64[j] - P64[i]
            dist = np.linalg.norm(v)
            if i != j and 0 < dist < 2.0:
                f += _circle_k(dist) * v / dist
                cnt += 1
        exp.append(f / max(cnt, 1))
    exp = np.array(exp)
    np.testing.assert_allclose(np.stack([X, Y, Z], 1), P64 + exp, rtol=0, atol=3e-6)
    np.testing.assert_allclose(D, np.linalg.norm(exp, axis=1), rtol=1e-5, atol=1e-9)
    assert D[3] == 0.0 and X[3] == np.float32(8.5)
    # the averaged force on A: B pushes (d=0.6 < 1), C pushes (d=0.8 < 1), half each
    assert exp[0][0] < 0 and exp[0][1] < 0


def test_circles_clamped_to_box():
    # pushed out through the low face: the position is clamped to 0 (the box is [0, max]^3)
    params = np.array([10.0, 2.0, 0.05], np.float32)
    s = O.Sim(0, params, 1)
    x = np.array([0.02, 0.32], np.float32)
    c = np.array([5.0, 5.0], np.float32)
    s.set_agents(0, [x, c.copy(), c.copy(), np.zeros(2, np.float32)])
    s.step(1)
    (X, _, _, D), _ = _agents_by_id(s, 0, [np.float32] * 4)
    k = _circle_k(0.3)
    assert 0.02 + k < 0 and X[0] == 0.0
    np.testing.assert_allclose(X[1], 0.32 - k, atol=2e-6)


BOIDS_P = None


def _boids_params():
    from flamegpu2_amd import workloads
    p, _ = workloads.boids(4)
    return p   # [min 0, max 3.2, r 0.05, sep 0.015, dt 1, global 0.15, steer 0.005, coll 0.01, match 0.125, vmin 0.025, vmax 0.05]


def _boids_run(P, V, steps=1):
    params = _boids_params()
    s = O.Sim(1, params, 7)
    P = np.asarray(P, np.float32)
    V = np.asarray(V, np.float32)
    s.set_agents(0, [P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy(), V[:, 0].copy(), V[:, 1].copy(), V[:, 2].copy()])
    s.step(steps)
    a, ids = _agents_by_id(s, 0, [np.float32] * 6)
    return np.stack(a[:3], 1), np.stack(a[3:], 1)


def _boids_expect(P, V, W=3.2):
    """One boids step by hand, float64, minimum-image displacement across the wrap."""
    G, STEER, COLL, MATCH, SEP, R = 0.15, 0.005, 0.01, 0.125, 0.015, 0.05
    VMIN, VMAX = 0.025, 0.05
    P = np.asarray(P, np.float32).astype(np.float64)
    V = np.asarray(V, np.float32).astype(np.float64)
    outP, outV = [], []
    for i in range(len(P)):
        near = []
        for j in range(len(P)):
            if j == i:
                continue
            d = P[j] - P[i]
            d -= W * np.round(d / W)          # position of j seen from i, across the wrap
            if np.linalg.norm(d) < R:
                near.append((d, V[j]))
        steer = np.zeros(3)
        if near:
            centre = np.mean([d for d, _ in near], axis=0)                 # cohesion: towards the centre
            steer += centre * STEER
            steer += np.mean([v for _, v in near], axis=0) * MATCH - V[i]  # alignment
            for d, _ in near:                                              # separation: away, (1 - d/SEP)^2
                dist = np.linalg.norm(d)
                if dist < SEP:
                    steer -= COLL * d * (1.0 - dist / SEP) ** 2
        v = V[i] + G * steer
        sp = np.linalg.norm(v)
        v = v / sp * VMAX if sp > VMAX else (v / sp * VMIN if sp < VMIN else v)
        outV.append(v)
        outP.append(np.mod(P[i] + v, W))
    return np.array(outP), np.array(outV)


@pytest.mark.parametrize("s,sign", [(0.003, -1.0), (0.006, +1.0)])
def test_boids_separation_against_cohesion(s, sign):
    # Two boids at rest, s apart on x: cohesion pulls with STEER*s, separation pushes with
    # COLL*s*(1 - s/SEP)^2; they balance where (1 - s/SEP)^2 = STEER/COLL, s* = SEP*(1 - 1/sqrt 2)
    # = 0.00439.  The resulting speed is tiny, so the minimum-speed clamp sets |v| = VMIN:
    # below s* A moves away from B (-x), above it towards B (+x).
    P = [[1.6, 1.6, 1.6], [1.6 + s, 1.6, 1.6]]
    V = [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]
    gotP, gotV = _boids_run(P, V)
    np.testing.assert_allclose(gotV[0], [sign * 0.025, 0, 0], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(gotV[1], [-sign * 0.025, 0, 0], rtol=1e-6, atol=1e-9)
    eP, eV = _boids_expect(P, V)
    np.testing.assert_allclose(gotV, eV, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(gotP, eP, rtol=0, atol=1e-6)


def test_boids_alignment_and_cohesion_no_separation():
    # 0.03 apart (outside SEP): A steers towards B's place and towards MATCH x B's velocity,
    # and away from its own velocity.  Speeds stay inside [VMIN, VMAX].
    P = [[1.0, 1.0, 1.0], [1.03, 1.0, 1.0]]
    V = [[0.04, 0.0, 0.0], [0.0, 0.04, 0.01]]
    gotP, gotV = _boids_run(P, V)
    eP, eV = _boids_expect(P, V)
    dv = gotV - np.asarray(V, np.float32).astype(np.float64)
    edv = eV - np.asarray(V, np.float32).astype(np.float64)
    np.testing.assert_allclose(dv, edv, rtol=1e-5, atol=1e-9)
    # by hand for A: 0.15 * ((0.03*0.005, 0, 0) + 0.125*(0, 0.04, 0.01) - (0.04, 0, 0))
    np.testing.assert_allclose(edv[0], 0.15 * np.array([0.03 * 0.005 - 0.04, 0.005, 0.00125]), rtol=1e-5)
    np.testing.assert_allclose(gotP, eP, rtol=0, atol=1e-6)


def test_boids_three_body_means():
    # A sees B and C (both within r, C within SEP): cohesion and alignment use the mean of the two
    P = [[2.0, 2.0, 2.0], [2.02, 2.01, 2.0], [2.0, 1.99, 2.008]]
    V = [[0.03, 0.0, 0.0], [0.0, 0.03, 0.0], [0.0, 0.0, -0.03]]
    gotP, gotV = _boids_run(P, V)
    eP, eV = _boids_expect(P, V)
    dv = gotV - np.asarray(V, np.float32).astype(np.float64)
    edv = eV - np.asarray(V, np.float32).astype(np.float64)
    np.testing.assert_allclose(dv, edv, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(gotP, eP, rtol=0, atol=1e-6)


def test_boids_speed_clamps():
    # alone: no steering; 0.1 is clamped down to VMAX, 0.02 and 0.01 raised to VMIN, 0.04 kept,
    # direction kept
    P = [[0.5, 0.5, 0.5], [2.5, 2.5, 2.5], [1.5, 0.5, 2.5], [0.5, 2.5, 1.5]]
    V = [[0.06, 0.08, 0.0], [0.0, 0.012, -0.016], [0.006, 0.0, 0.008], [0.024, 0.0, -0.032]]
    gotP, gotV = _boids_run(P, V)
    np.testing.assert_allclose(gotV[0], [0.03, 0.04, 0.0], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(gotV[1], [0.0, 0.015, -0.02], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(gotV[2], [0.015, 0.0, 0.02], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(gotV[3], np.asarray(V[3], np.float32), rtol=0, atol=0)
    np.testing.assert_allclose(gotP, np.asarray(P) + gotV, rtol=0, atol=1e-6)


def test_boids_pair_across_the_wrap():
    # A pair straddling x = 0 / 3.2 interacts through the wrap exactly as the same pair away
    # from the boundary: velocities agree, and A's position wraps to the far side.
    Pw = [[0.002, 1.0, 1.0], [3.196, 1.0, 1.0]]     # 0.006 apart across the wrap
    Pi = [[1.602, 1.0, 1.0], [1.596, 1.0, 1.0]]     # the same pair, shifted by 1.6
    V = [[-0.03, 0.0, 0.0], [0.0, 0.03, 0.0]]
    gw, vw = _boids_run(Pw, V)
    gi, vi = _boids_run(Pi, V)
    np.testing.assert_allclose(vw, vi, rtol=1e-4, atol=1e-8)
    eP, eV = _boids_expect(Pw, V)
    np.testing.assert_allclose(vw, eV, rtol=1e-4, atol=1e-8)
    assert gw[0][0] > 3.0   # moved from 0.002 by about -0.026: wrapped to about 3.176
    np.testing.assert_allclose(gw, eP, rtol=0, atol=1e-5)


def _pp_params(width=1000.0):
    from flamegpu2_amd import workloads
    p, _, _ = workloads.predprey(1, 1, seed=3, width=width)
    return p   # [width, radius 5, eat_r 1, p_prey_death .005, p_prey_repro .02, p_pred_repro .01, starve 50, jitter 1, flee .5, hunt .5]


def _uniform(seed, aid, step, stream, k):
    """The keyed agent stream from the published Philox4x32-10 (pinned above) and the
    reference's float mapping (AgentRandom.cuh:80-83)."""
    blk = O.philox([k >> 2, stream, step, aid], [seed & 0xFFFFFFFF, seed >> 32])
    return float(np.float32((blk[k & 3] & 0xFFFFFF) / 16777216.0))


def _pp_sim(seed, prey, pred, width=1000.0):
    s = O.Sim(2, _pp_params(width), seed)
    s.set_agents(0, [np.asarray(a, np.float32) for a in prey])
    s.set_agents(1, [np.asarray(a, np.float32) for a in pred[:4]] + [np.asarray(pred[4], np.uint32)])
    return s


@pytest.mark.parametrize("birth", [True, False])
def test_predprey_predator_eats_and_reproduces(birth):
    # One prey 0.5 from one predator: the prey is eaten whatever its draws; the predator's
    # hunger resets, it turns towards the prey (hunt 0.5, speed capped at 1), moves, and
    # reproduces iff its draw (stream 7, draw 0) is below 0.01.  The child sits at the
    # parent's new place plus (u - 0.5) * jitter, with velocity 2u - 1, hunger 0, ID 2.
    seed = next(sd for sd in range(1, 20000) if (_uniform(sd, 1, 0, 7, 0) < 0.01) == birth)
    prey = [[500.0], [500.0], [0.3], [0.0]]
    pred = [[500.5], [500.0], [0.2], [0.4], [7]]
    s = _pp_sim(seed, prey, pred)
    s.step(1)
    assert s.count(0) == 0
    (PX, PY, PVX, PVY, LIFE), ids = _agents_by_id(s, 1, [np.float32] * 4 + [np.uint32])
    vx, vy = 0.2 + (-0.5 / 0.5) * 0.5, 0.4        # nearest prey at dx = -0.5
    sp = np.hypot(vx, vy)
    if sp > 1.0:
        vx, vy = vx / sp, vy / sp
    px, py = 500.5 + vx, 500.0 + vy
    assert ids[0] == 1
    np.testing.assert_allclose([PX[0], PY[0], PVX[0], PVY[0]], [px, py, vx, vy], atol=1e-4)
    assert LIFE[0] == 0
    if birth:
        assert list(ids) == [1, 2]
        u = [_uniform(seed, 1, 0, 7, k) for k in range(5)]
        np.testing.assert_allclose([PX[1], PY[1], PVX[1], PVY[1]],
                                   [px + (u[1] - 0.5), py + (u[2] - 0.5), 2 * u[3] - 1, 2 * u[4] - 1], atol=1e-4)
        assert LIFE[1] == 0
    else:
        assert list(ids) == [1]


@pytest.mark.parametrize("life0", [0, 13, 49])
def test_predprey_predator_starves_after_50_hungry_steps(life0):
    # no prey: hunger rises by one per step and the predator dies at the end of the step in
    # which it reaches 50; it never reproduces (it has not eaten)
    s = _pp_sim(5, [[], [], [], []], [[100.0], [100.0], [0.5], [0.5], [life0]])
    left = 50 - life0
    s.step(left - 1)
    (_, _, _, _, LIFE), ids = _agents_by_id(s, 1, [np.float32] * 4 + [np.uint32])
    assert list(ids) == [1] and LIFE[0] == 49
    s.step(1)
    assert s.count(1) == 0


def test_predprey_prey_alone_follow_their_draws():
    # 40 prey, no predators, 25 steps: each prey dies when its stream-1 draw is below 0.005 and
    # otherwise gives birth when its stream-3 draw is below 0.02; children get the next IDs in
    # parent order (here parent order = list order after the spatial sort, so compare sets and
    # counts) and their values come from the parent's draws.  Speeds above 1 are capped at 1.
    seed, n = 11, 40
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1000, n).astype(np.float32)
    y = rng.uniform(0, 1000, n).astype(np.float32)
    vx = rng.uniform(-0.6, 0.6, n).astype(np.float32)
    vy = rng.uniform(-0.6, 0.6, n).astype(np.float32)
    s = _pp_sim(seed, [x, y, vx, vy], [[], [], [], [], []])
    alive = {i + 1: (x[i], y[i], vx[i], vy[i]) for i in range(n)}
    next_id = n + 1
    W = np.float32(1000.0)

    def wrap(v):
        v = np.float32(v)
        if v < 0:
            v = np.float32(v + W)
        if v >= W:
            v = np.float32(v - W)
        return v
    for step in range(25):
        s.step(1)
        survivors = {}
        for aid, (px, py, pvx, pvy) in alive.items():
            if _uniform(seed, aid, step, 1, 0) < 0.005:
                continue
            sp = np.sqrt(np.float32(pvx * pvx + pvy * pvy), dtype=np.float32)
            if sp > 1.0:   # the flee function caps the speed at 1 (children start at up to sqrt 2)
                pvx, pvy = np.float32(pvx / sp), np.float32(pvy / sp)
            survivors[aid] = (wrap(px + pvx), wrap(py + pvy), pvx, pvy)
        got = s.get_agents(0, [np.float32] * 4)
        (GX, GY, GVX, GVY), gids = got
        parents = [aid for aid in gids if aid in survivors]   # parents in list order at birth time
        born = {}
        for aid in parents:
            if _uniform(seed, aid, step, 3, 0) < 0.02:
                px, py = survivors[aid][:2]
                u = [_uniform(seed, aid, step, 3, k) for k in range(5)]
                born[next_id] = (wrap(np.float32(px) + np.float32((np.float32(u[1]) - np.float32(0.5)) * np.float32(1.0))),
                                 wrap(np.float32(py) + np.float32((np.float32(u[2]) - np.float32(0.5)) * np.float32(1.0))),
                                 np.float32(np.float32(u[3]) * np.float32(2.0) - np.float32(1.0)),
                                 np.float32(np.float32(u[4]) * np.float32(2.0) - np.float32(1.0)))
                next_id += 1
        alive = {**survivors, **born}
        assert sorted(gids.tolist()) == sorted(alive), f"step {step}"
        for i, aid in enumerate(gids):
            np.testing.assert_allclose([GX[i], GY[i], GVX[i], GVY[i]], alive[aid], rtol=0, atol=1e-4)
    assert next_id > n + 1 and len(alive) != n   # both births and deaths happened
