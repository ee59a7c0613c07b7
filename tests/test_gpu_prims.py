"""GPU parity of the libfgx primitives against the serial oracle (bit-exact)."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, empty, host, ptr, sync, scratch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from flamegpu2_amd import fgx
    c = fgx.Context(0)
    yield c
    c.close()


def _build_case(ctx, dims, n, width, radius, seed, use_dn=False, spread=0.0):
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(seed)
    P = rng.uniform(-spread * width, (1 + spread) * width, size=(dims, n)).astype(np.float32)
    ids = (np.arange(n, dtype=np.uint32) * 7 + 3).astype(np.uint32)
    extra = rng.uniform(size=n)                          # 8-byte variable
    mn, mx = [0.0] * dims, [float(width)] * dims
    mo = O.spatial_meta(dims, mn, mx, radius)
    pbm_o, perm_o, outs_o = O.spatial_build(mo, P[0], P[1], P[2] if dims == 3 else None, extra=[ids, extra])
    m = fgx.spatial_meta(dims, mn, mx, radius)
    assert m.bins == mo.bins
    d_in = [dev(P[a]) for a in range(dims)] + [dev(ids), dev(extra)]
    d_out = [empty(n, np.float32) for _ in range(dims)] + [empty(n, np.uint32), empty(n, np.float64)]
    nbytes = [4] * dims + [4, 8]
    pbm = empty(m.bins + 1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, m.bins)
    s = scratch(sb)
    d_n = None
    cap = n
    if use_dn:
        cnt = dev(np.array([n], np.uint32))
        d_n = ptr(cnt)
        cap = n  # capacity == n; the count is read on the device
    for _ in range(2):   # the second call reuses the scratch as the first left it
        fgx.spatial_build(ctx, m, cap, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]) if dims == 3 else None,
                          [ptr(t) for t in d_in], [ptr(t) for t in d_out], nbytes, ptr(pbm), ptr(s), sb, d_n=d_n)
    sync()
    assert np.array_equal(host(pbm, np.uint32, m.bins + 1), pbm_o)
    for a in range(dims):
        assert np.array_equal(host(d_out[a], np.float32, n), outs_o[a])
    assert np.array_equal(host(d_out[dims], np.uint32, n), outs_o[dims])
    assert np.array_equal(host(d_out[dims + 1], np.float64, n), outs_o[dims + 1])


@pytest.mark.parametrize("dims,n,width,radius,seed", [
    (3, 16384, float(np.float32(np.cbrt(16384.0))), 2.0, 12),     # circles config
    (3, 1 << 20, 3.2, 0.05, 7),                                   # boids config: 64^3 bins
    (2, 250000, 1000.0, 5.0, 3),                                  # predator-prey grid
    (3, 1, 5.0, 1.0, 1),
    (3, 2049, 5.0, 1.0, 2),
])
def test_spatial_build_bit_exact(ctx, dims, n, width, radius, seed):
    _build_case(ctx, dims, n, width, radius, seed)


def test_spatial_build_out_of_range_and_device_count(ctx):
    _build_case(ctx, 3, 50000, 10.0, 1.0, 5, use_dn=True, spread=0.3)


def test_spatial_build_empty(ctx):
    from flamegpu2_amd import fgx
    m = fgx.spatial_meta(3, [0, 0, 0], [4, 4, 4], 1.0)
    pbm = dev(np.full(m.bins + 1, 7, np.uint32))
    sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, 0, m.bins)
    s = scratch(sb)
    fgx.spatial_build(ctx, m, 0, None, None, None, [], [], [], ptr(pbm), ptr(s), sb)
    sync()
    assert np.all(host(pbm, np.uint32) == 0)


def test_spatial_build_clumped_bins(ctx):
    # many messages in a few bins: the stable in-bin ranking must still hold
    rng = np.random.default_rng(9)
    n = 20000
    P = rng.uniform(0, 0.3, size=(3, n)).astype(np.float32)   # 27 bins of a 10^3 grid
    from flamegpu2_amd import fgx
    mo = O.spatial_meta(3, [0] * 3, [1.0] * 3, 0.1)
    ids = np.arange(n, dtype=np.uint32)
    pbm_o, perm_o, outs_o = O.spatial_build(mo, P[0], P[1], P[2], extra=[ids])
    m = fgx.spatial_meta(3, [0] * 3, [1.0] * 3, 0.1)
    d_in = [dev(P[a]) for a in range(3)] + [dev(ids)]
    d_out = [empty(n, np.float32) for _ in range(3)] + [empty(n, np.uint32)]
    pbm = empty(m.bins + 1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, m.bins)
    s = scratch(sb)
    fgx.spatial_build(ctx, m, n, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]), [ptr(t) for t in d_in],
                      [ptr(t) for t in d_out], [4, 4, 4, 4], ptr(pbm), ptr(s), sb)
    sync()
    assert np.array_equal(host(pbm, np.uint32, m.bins + 1), pbm_o)
    assert np.array_equal(host(d_out[3], np.uint32, n), outs_o[3])


def _timed(fn, reps=3):
    """Best wall time of `reps` synchronised calls, after one untimed call."""
    import time
    fn()
    sync()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        sync()
        best = min(best, time.perf_counter() - t0)
    return best


def _spatial_build_check(ctx, P, mn, mx, radius, max_seconds):
    """Builds the message list of positions P (3 x n) twice (scratch reused), checks the bin
    table and the stable message order bit for bit against the oracle, and bounds the time."""
    from flamegpu2_amd import fgx
    n = P.shape[1]
    ids = np.arange(n, dtype=np.uint32)
    mo = O.spatial_meta(3, mn, mx, radius)
    pbm_o, perm_o, outs_o = O.spatial_build(mo, P[0], P[1], P[2], extra=[ids])
    m = fgx.spatial_meta(3, mn, mx, radius)
    d_in = [dev(P[a]) for a in range(3)] + [dev(ids)]
    d_out = [empty(n, np.float32) for _ in range(3)] + [empty(n, np.uint32)]
    pbm = empty(m.bins + 1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, m.bins)
    s = scratch(sb)

    def run():
        fgx.spatial_build(ctx, m, n, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]), [ptr(t) for t in d_in],
                          [ptr(t) for t in d_out], [4, 4, 4, 4], ptr(pbm), ptr(s), sb)
    t = _timed(run)
    assert np.array_equal(host(pbm, np.uint32, m.bins + 1), pbm_o)
    assert np.array_equal(host(d_out[3], np.uint32, n), outs_o[3])
    for a in range(3):
        assert np.array_equal(host(d_out[a], np.float32, n), outs_o[a])
    print(f"spatial build n={n} bins={m.bins} largest bin={int(np.diff(pbm_o).max())}: {t * 1e3:.3f} ms")
    assert t < max_seconds, (t, max_seconds)
    return pbm_o


def test_spatial_build_1M_in_64_bins(ctx):
    # 2^20 messages in random order over a 4^3 grid: every bin holds ~16 K members, so the
    # in-bin order comes from the per-bin bitmap ranks (a quadratic rank would take seconds)
    P = np.random.default_rng(64).uniform(0, 4.0, size=(3, 1 << 20)).astype(np.float32)
    pbm = _spatial_build_check(ctx, P, [0.0] * 3, [4.0] * 3, 1.0, 0.05)
    assert len(pbm) == 65


def test_spatial_build_bins_of_every_size(ctx):
    # a 64^3 grid with 1.5 M messages: one bin holds 700 K of them spread over the whole
    # input (three bitmap windows), 100 bins hold 200 each (LDS ranks), 40 bins hold 1,000
    # each (one bitmap window) and the rest are uniform (at most a few per bin)
    rng = np.random.default_rng(11)
    n = 1_500_000
    P = rng.uniform(0, 3.2, size=(3, n)).astype(np.float32)
    idx = rng.permutation(n)
    giant, k = idx[:700_000], 700_000
    P[:, giant] = np.float32(1.6125)
    cell = np.float32(0.05)
    for j in range(100):
        sel = idx[k:k + 200]
        k += 200
        c = rng.integers(0, 64, size=3).astype(np.float32)
        P[:, sel] = ((c[:, None] + 0.25 + 0.5 * rng.uniform(size=(1, 200))) * cell).astype(np.float32)
    for j in range(40):
        sel = idx[k:k + 1000]
        k += 1000
        c = rng.integers(0, 64, size=3).astype(np.float32)
        P[:, sel] = ((c[:, None] + 0.25 + 0.5 * rng.uniform(size=(3, 1000))) * cell).astype(np.float32)
    pbm = _spatial_build_check(ctx, P, [0.0] * 3, [3.2] * 3, 0.05, 0.05)
    sizes = np.diff(pbm)
    assert sizes.max() >= 700_000 and ((sizes > 32) & (sizes <= 512)).sum() >= 50 and ((sizes > 512) & (sizes < 5000)).sum() >= 30


def test_spatial_build_one_bin(ctx):
    # every message in the same bin
    P = np.full((3, 300_000), 0.5, np.float32)
    _spatial_build_check(ctx, P, [0.0] * 3, [3.2] * 3, 0.05, 0.05)


@pytest.mark.parametrize("n,keys", [(1 << 20, 8), (100_000, 1), (70_000, 3000)])
def test_bucket_build_few_keys(ctx, n, keys):
    # bucket index (a stable sort by integer key) with few keys, random order
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(n + keys)
    key_min = -3
    key = rng.integers(key_min, key_min + keys, size=n).astype(np.int32)
    val = rng.uniform(size=n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    perm = O.stable_sort_perm((key - key_min).astype(np.uint32))
    table = np.zeros(keys + 1, np.uint32)
    table[1:] = np.cumsum(np.bincount(key - key_min, minlength=keys)).astype(np.uint32)
    d_key, d_val, d_ids = dev(key), dev(val), dev(ids)
    o_key, o_val, o_ids = empty(n, np.int32), empty(n, np.float32), empty(n, np.uint32)
    pbm = empty(keys + 1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, keys)
    s = scratch(sb)

    def run():
        fgx.bucket_build(ctx, key_min, keys, n, ptr(d_key), [ptr(d_key), ptr(d_val), ptr(d_ids)],
                         [ptr(o_key), ptr(o_val), ptr(o_ids)], [4, 4, 4], ptr(pbm), ptr(s), sb)
    t = _timed(run)
    assert np.array_equal(host(pbm, np.uint32, keys + 1), table)
    assert np.array_equal(host(o_ids, np.uint32, n), perm)
    assert np.array_equal(host(o_key, np.int32, n), key[perm])
    assert np.array_equal(host(o_val, np.float32, n), val[perm])
    print(f"bucket build n={n} keys={keys}: {t * 1e3:.3f} ms")
    assert t < 0.05, t


@pytest.mark.parametrize("dims,n,mn,mx,r,seed", [
    (3, 1 << 20, [0] * 3, [3.2] * 3, 0.05, 7),
    (3, 16384, [0] * 3, [25.398416] * 3, 2.0, 12),
    (2, 100000, [0, 0], [1000, 1000], 5.0, 3),
    (3, 4, [-5] * 3, [5] * 3, 0.2, 0),
])
def test_agent_sort_bit_exact(ctx, dims, n, mn, mx, r, seed):
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(seed)
    lo, hi = np.float32(mn[0]), np.float32(mx[0])
    P = rng.uniform(lo - 0.1 * (hi - lo), hi + 0.1 * (hi - lo), size=(dims, n)).astype(np.float32)
    if n == 4:
        P = np.stack([-np.arange(4, dtype=np.float32)] * dims)
    keys, bits = O.agent_sort_keys(dims, mn, mx, r, P[0], P[1], P[2] if dims == 3 else None)
    perm = O.stable_sort_perm(keys)
    order = np.arange(n, dtype=np.uint32)
    d_in = [dev(P[a]) for a in range(dims)] + [dev(order)]
    d_out = [empty(n, np.float32) for _ in range(dims)] + [empty(n, np.uint32)]
    sb = fgx.scratch_bytes(fgx.FGX_OP_AGENT_SORT, n, 1 << bits)
    s = scratch(sb)
    fgx.agent_sort(ctx, dims, mn, mx, r, n, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]) if dims == 3 else None,
                   [ptr(t) for t in d_in], [ptr(t) for t in d_out], [4] * (dims + 1), ptr(s), sb)
    sync()
    assert np.array_equal(host(d_out[dims], np.uint32, n), perm)
    if n == 4:
        assert perm.tolist() == [3, 2, 1, 0]


@pytest.mark.parametrize("n,p,keep", [(1, 1.0, 1), (2047, 0.5, 1), (2048, 0.9, 1), (2049, 0.1, 0),
                                      (1_250_000, 0.7, 1), (0, 0.5, 1),
                                      (3_000_000, 0.5, 1)])   # > 1,024 tiles: the tile counts are block-scanned
def test_compact_stable(ctx, n, p, keep):
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(n)
    flags = (rng.uniform(size=n) < p).astype(np.uint32)
    a = rng.uniform(size=n).astype(np.float32)
    b = np.arange(n, dtype=np.uint32)
    c = rng.uniform(size=n)
    cnt_o, outs_o = O.compact(flags, keep, [a, b, c])
    d_f = dev(flags)
    d_in = [dev(a), dev(b), dev(c)]
    d_out = [empty(n, np.float32), empty(n, np.uint32), empty(n, np.float64)]
    d_cnt = empty(1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_COMPACT, n)
    s = scratch(sb)
    for _ in range(3):
        fgx.compact(ctx, ptr(d_f), keep, n, [ptr(t) for t in d_in], [ptr(t) for t in d_out], [4, 4, 8],
                    ptr(d_cnt), ptr(s), sb)
    sync()
    cnt = int(host(d_cnt, np.uint32)[0])
    assert cnt == cnt_o
    assert np.array_equal(host(d_out[0], np.float32, cnt), outs_o[0])
    assert np.array_equal(host(d_out[1], np.uint32, cnt), outs_o[1])
    assert np.array_equal(host(d_out[2], np.float64, cnt), outs_o[2])


@pytest.mark.parametrize("dims,n,width,radius,seed", [(3, 200_000, 10.0, 0.5, 31), (2, 50_000, 30.0, 1.0, 32)])
def test_spatial_build_records(ctx, dims, n, width, radius, seed):
    # the build's scatter writes the {x, y, z, id} records of the sorted list: word k of
    # record i is variable rec_vars[k] of sorted element i (z absent in 2-D: 0 words)
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(seed)
    P = rng.uniform(0, width, size=(dims, n)).astype(np.float32)
    ids = (np.arange(n, dtype=np.uint32) * 5 + 1).astype(np.uint32)
    mn, mx = [0.0] * dims, [float(width)] * dims
    pbm_o, _, outs_o = O.spatial_build(O.spatial_meta(dims, mn, mx, radius), P[0], P[1], P[2] if dims == 3 else None,
                                       extra=[ids])
    m = fgx.spatial_meta(dims, mn, mx, radius)
    d_in = [dev(P[a]) for a in range(dims)] + [dev(ids)]
    d_out = [empty(n, np.float32) for _ in range(dims)] + [empty(n, np.uint32)]
    rec = empty(4 * n, np.uint32)
    pbm = empty(m.bins + 1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, m.bins)
    s = scratch(sb)
    rec_vars = [0, 1, 2, 3] if dims == 3 else [0, 1, -1, 2]
    fgx.spatial_build_records(ctx, m, n, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]) if dims == 3 else None,
                              [ptr(t) for t in d_in], [ptr(t) for t in d_out], [4] * (dims + 1), rec_vars, ptr(rec),
                              ptr(pbm), ptr(s), sb)
    sync()
    assert np.array_equal(host(pbm, np.uint32, m.bins + 1), pbm_o)
    got = host(rec, np.uint32, 4 * n).reshape(n, 4)
    for a in range(dims):
        assert np.array_equal(host(d_out[a], np.float32, n), outs_o[a])
        assert np.array_equal(got[:, a].view(np.float32), outs_o[a])
    if dims == 2:
        assert not got[:, 2].any()
    assert np.array_equal(got[:, 3], outs_o[dims])
    with pytest.raises(Exception):   # a record word must name a 4-byte variable
        fgx.spatial_build_records(ctx, m, n, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]) if dims == 3 else None,
                                  [ptr(t) for t in d_in], [ptr(t) for t in d_out], [4] * (dims + 1), [0, 1, 7, -1],
                                  ptr(rec), ptr(pbm), ptr(s), sb)


@pytest.mark.parametrize("use_dn", [False, True])
def test_pack_records(ctx, use_dn):
    # the neighbour walk's {x, y, z, id} records: word k of record i is source k's element i
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(11)
    cap, n = 100_003, 77_777
    w = [rng.integers(0, 2**32, size=cap, dtype=np.uint32) for _ in range(4)]
    d_w = [dev(a) for a in w]
    out = empty(4 * cap, np.uint32)
    d_n = dev(np.array([n], np.uint32)) if use_dn else None
    fgx.pack_records(ptr(d_n) if use_dn else None, cap, ptr(d_w[0]), ptr(d_w[1]), None, ptr(d_w[3]), ptr(out))
    sync()
    got = host(out, np.uint32, 4 * cap).reshape(cap, 4)
    m = n if use_dn else cap
    assert np.array_equal(got[:m, 0], w[0][:m]) and np.array_equal(got[:m, 1], w[1][:m])
    assert not got[:m, 2].any()                      # NULL source: 0 words
    assert np.array_equal(got[:m, 3], w[3][:m])
    assert not got[m:].any()                         # records past the count are not written
    with pytest.raises(Exception):
        fgx.pack_records(None, cap, ptr(d_w[0]), None, None, None, ptr(out) + 4)   # not 16-byte aligned


def test_compact_device_count(ctx):
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(3)
    cap, n = 100000, 61234
    flags = (rng.uniform(size=cap) < 0.4).astype(np.uint32)
    b = np.arange(cap, dtype=np.uint32)
    cnt_o, outs_o = O.compact(flags[:n], 1, [b[:n]])
    d_f, d_b, d_o = dev(flags), dev(b), empty(cap, np.uint32)
    d_n = dev(np.array([n], np.uint32))
    d_cnt = empty(1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_COMPACT, cap)
    s = scratch(sb)
    fgx.compact(ctx, ptr(d_f), 1, cap, [ptr(d_b)], [ptr(d_o)], [4], ptr(d_cnt), ptr(s), sb, d_n=ptr(d_n))
    sync()
    cnt = int(host(d_cnt, np.uint32)[0])
    assert cnt == cnt_o and np.array_equal(host(d_o, np.uint32, cnt), outs_o[0])


def test_compact_count_in_place(ctx):
    # the output count is the input count (a list compacted in place): the second pass
    # must not see the count its own last tile writes
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(5)
    for cap, n in [(5_000_000, 4_500_001), (300_000, 299_999)]:
        flags = (rng.uniform(size=cap) < 0.3).astype(np.uint32)
        b = np.arange(cap, dtype=np.uint32)
        cnt_o, outs_o = O.compact(flags[:n], 1, [b[:n]])
        d_f, d_b, d_o = dev(flags), dev(b), empty(cap, np.uint32)
        d_n = dev(np.array([n], np.uint32))
        sb = fgx.scratch_bytes(fgx.FGX_OP_COMPACT, cap)
        s = scratch(sb)
        fgx.compact(ctx, ptr(d_f), 1, cap, [ptr(d_b)], [ptr(d_o)], [4], ptr(d_n), ptr(s), sb, d_n=ptr(d_n))
        sync()
        cnt = int(host(d_n, np.uint32)[0])
        assert cnt == cnt_o and np.array_equal(host(d_o, np.uint32, cnt), outs_o[0])


def test_spatial_build_keep_scratch_zero():
    """FGX_OPT_KEEP_SCRATCH_ZERO: sorts leave their histogram zeroed and skip the clearing
    launch when a buffer is reused with the same bin count; a different bin count on the
    same buffer clears again.  Every build must still equal the oracle's."""
    from flamegpu2_amd import fgx
    c = fgx.Context(0)
    c.set_option(fgx.FGX_OPT_KEEP_SCRATCH_ZERO, 1)
    rng = np.random.default_rng(17)
    n = 50000
    sb = max(fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, 40 ** 3), fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, 10 ** 3))
    s = scratch(sb)   # zeroed on allocation
    for width, radius in [(4.0, 0.1), (4.0, 0.1), (4.0, 0.4), (4.0, 0.1), (4.0, 0.1)]:
        P = rng.uniform(0, width, size=(3, n)).astype(np.float32)
        mo = O.spatial_meta(3, [0.0] * 3, [width] * 3, radius)
        pbm_o, perm_o, outs_o = O.spatial_build(mo, P[0], P[1], P[2])
        m = fgx.spatial_meta(3, [0.0] * 3, [width] * 3, radius)
        d_in = [dev(P[a]) for a in range(3)]
        d_out = [empty(n, np.float32) for _ in range(3)]
        pbm = empty(m.bins + 1, np.uint32)
        fgx.spatial_build(c, m, n, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]), [ptr(t) for t in d_in],
                          [ptr(t) for t in d_out], [4, 4, 4], ptr(pbm), ptr(s), sb)
        sync()
        assert np.array_equal(host(pbm, np.uint32, m.bins + 1), pbm_o), (width, radius)
        for a in range(3):
            assert np.array_equal(host(d_out[a], np.float32, n), outs_o[a])
    c.close()


def test_spatial_build_captured_on_fresh_scratch():
    """A build captured as the first sort of a freshly cleared buffer (FGX_OPT_KEEP_SCRATCH_ZERO)
    carries no clearing launch; every replay, with eager builds of the same bin count on the
    same buffer between the replays, still equals the oracle."""
    from flamegpu2_amd import fgx
    c = fgx.Context(0)
    c.set_option(fgx.FGX_OPT_KEEP_SCRATCH_ZERO, 1)
    rng = np.random.default_rng(23)
    n, width, radius = 60000, 4.0, 0.1
    m = fgx.spatial_meta(3, [0.0] * 3, [width] * 3, radius)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, n, m.bins)
    s = scratch(sb)   # cleared by fgx_scratch_clear, no sort yet
    cases = []
    for _ in range(2):
        P = rng.uniform(0, width, size=(3, n)).astype(np.float32)
        pbm_o, _, outs_o = O.spatial_build(O.spatial_meta(3, [0.0] * 3, [width] * 3, radius), P[0], P[1], P[2])
        d_in = [dev(P[a]) for a in range(3)]
        d_out = [empty(n, np.float32) for _ in range(3)]
        cases.append((d_in, d_out, empty(m.bins + 1, np.uint32), pbm_o, outs_o))
    st = fgx.stream_create()

    def build(case):
        d_in, d_out, pbm = case[0], case[1], case[2]
        fgx.spatial_build(c, m, n, ptr(d_in[0]), ptr(d_in[1]), ptr(d_in[2]), [ptr(t) for t in d_in],
                          [ptr(t) for t in d_out], [4, 4, 4], ptr(pbm), ptr(s), sb, stream=st)

    def check(case):
        fgx.stream_synchronize(st)
        assert np.array_equal(host(case[2], np.uint32, m.bins + 1), case[3])
        for a in range(3):
            assert np.array_equal(host(case[1][a], np.float32, n), case[4][a])
        for t in case[1] + [case[2]]:
            fgx.fill_u32(ptr(t), 0, t.alloc_bytes // 4)   # so that each check sees fresh output
        sync()

    g = fgx.graph_capture(st, lambda: build(cases[0]))
    try:
        for _ in range(3):
            fgx.graph_launch(g, st)
            check(cases[0])
            build(cases[1])
            check(cases[1])
    finally:
        fgx.graph_destroy(g)
        fgx.stream_destroy(st)
    c.close()


def test_spatial_build_large_bin_table(ctx):
    # 4.1 M bins (> 1,024 scan tiles): the bin counts go through the block-scanned path
    _build_case(ctx, 3, 1 << 20, 16.0, 0.1, 21)


@pytest.mark.parametrize("n_par,existing,p", [(5000, 1234, 0.3), (1_000_000, 500_000, 0.02), (3, 0, 1.0),
                                             (3_000_000, 100, 0.01)])
def test_births_bit_exact(ctx, n_par, existing, p):
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(n_par)
    flags = (rng.uniform(size=n_par) < p).astype(np.uint32)
    newx = rng.uniform(size=n_par).astype(np.float32)
    cap = existing + n_par
    dstx = np.zeros(cap, np.float32); dstx[:existing] = 7.0
    dstid = np.zeros(cap, np.uint32); dstid[:existing] = np.arange(existing) + 1
    ox, oid = dstx.copy(), dstid.copy()
    c_o, nid_o = O.births(flags, [newx, None], [ox, oid], existing, 1, existing + 1)
    d_f, d_new = dev(flags), dev(newx)
    d_x, d_id = dev(dstx), dev(dstid)
    d_cnt = dev(np.array([existing], np.uint32))
    d_nid = dev(np.array([existing + 1], np.uint32))
    sb = fgx.scratch_bytes(fgx.FGX_OP_BIRTHS, n_par)
    s = scratch(sb)
    fgx.append_births(ctx, ptr(d_f), n_par, [ptr(d_new), 0], [ptr(d_x), ptr(d_id)], [4, 4], 1, ptr(d_cnt), cap,
                      ptr(d_nid), ptr(s), sb)
    sync()
    assert int(host(d_cnt, np.uint32)[0]) == c_o
    assert int(host(d_nid, np.uint32)[0]) == nid_o
    assert np.array_equal(host(d_x, np.float32, c_o), ox[:c_o])
    assert np.array_equal(host(d_id, np.uint32, c_o), oid[:c_o])


@pytest.mark.parametrize("n_par", [5000, 3_000_000])
def test_births_ids_by_parent_index(ctx, n_par):
    # the ID word set (a parent read its child's ID in-kernel): child ID = next ID + parent
    # index, the next ID advances by the parent count, and the append clears the word
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(n_par + 1)
    existing = 77
    flags = (rng.uniform(size=n_par) < 0.1).astype(np.uint32)
    newx = rng.uniform(size=n_par).astype(np.float32)
    cap = existing + n_par
    ox, oid = np.zeros(cap, np.float32), np.zeros(cap, np.uint32)
    c_o, nid_o = O.births(flags, [newx, None], [ox, oid], existing, 1, 1000, by_parent_index=True)
    d_f, d_new = dev(flags), dev(newx)
    d_x, d_id = empty(cap, np.float32), empty(cap, np.uint32)
    d_cnt = dev(np.array([existing], np.uint32))
    d_nid = dev(np.array([1000], np.uint32))
    d_slots = dev(np.array([1], np.uint32))
    sb = fgx.scratch_bytes(fgx.FGX_OP_BIRTHS, n_par)
    s = scratch(sb)
    fgx.append_births(ctx, ptr(d_f), n_par, [ptr(d_new), 0], [ptr(d_x), ptr(d_id)], [4, 4], 1, ptr(d_cnt), cap,
                      ptr(d_nid), ptr(s), sb, d_id_slots=ptr(d_slots))
    sync()
    assert int(host(d_cnt, np.uint32)[0]) == c_o
    assert int(host(d_nid, np.uint32)[0]) == nid_o == 1000 + n_par
    assert int(host(d_slots, np.uint32)[0]) == 0
    assert np.array_equal(host(d_x, np.float32, c_o)[existing:], ox[existing:c_o])
    assert np.array_equal(host(d_id, np.uint32, c_o)[existing:], oid[existing:c_o])
    # the word is clear: the next append numbers by rank again
    fgx.append_births(ctx, ptr(d_f), n_par, [ptr(d_new), 0], [ptr(d_x), ptr(d_id)], [4, 4], 1, ptr(d_cnt), cap,
                      ptr(d_nid), ptr(s), sb, d_id_slots=ptr(d_slots))
    sync()
    assert int(host(d_nid, np.uint32)[0]) == 1000 + n_par + int(flags.sum())


def test_births_overflow_flag(ctx):
    from flamegpu2_amd import fgx
    n_par, existing, cap = 1000, 900, 950
    flags = np.ones(n_par, np.uint32)
    d_f, d_new = dev(flags), dev(np.arange(n_par, dtype=np.float32))
    d_x = empty(cap, np.float32); d_id = empty(cap, np.uint32)
    d_cnt = dev(np.array([existing], np.uint32)); d_nid = dev(np.array([1], np.uint32))
    d_ovf = empty(1, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_BIRTHS, n_par)
    s = scratch(sb)
    fgx.append_births(ctx, ptr(d_f), n_par, [ptr(d_new), 0], [ptr(d_x), ptr(d_id)], [4, 4], 1, ptr(d_cnt), cap,
                      ptr(d_nid), ptr(s), sb, d_overflow=ptr(d_ovf))
    sync()
    assert int(host(d_cnt, np.uint32)[0]) == cap
    assert int(host(d_ovf, np.uint32)[0]) == 1
    assert np.array_equal(host(d_x, np.float32)[existing:cap], np.arange(cap - existing, dtype=np.float32))


@pytest.mark.parametrize("n", [1, 2, 2047, 2048, 2049, 262145, 10_000_000])
def test_exclusive_scan(ctx, n):
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(n)
    a = rng.integers(0, 64, size=n).astype(np.uint32)
    expect = np.concatenate([[0], np.cumsum(a.astype(np.uint64))[:-1]]).astype(np.uint32)
    d_a, d_o = dev(a), empty(n, np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SCAN, n)
    s = scratch(sb)
    fgx.exclusive_scan(ctx, ptr(d_a), ptr(d_o), n, ptr(s), sb)
    fgx.exclusive_scan(ctx, ptr(d_a), ptr(d_o), n, ptr(s), sb)
    sync()
    assert np.array_equal(host(d_o, np.uint32, n), expect)


def test_reduce_sum_deterministic(ctx):
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(1)
    n = 3_000_001
    a = rng.uniform(size=n).astype(np.float32)
    u = rng.integers(0, 1000, size=n).astype(np.uint32)
    sb = fgx.scratch_bytes(fgx.FGX_OP_SCAN, n)
    s = scratch(sb)
    d_a, d_u = dev(a), dev(u)
    outs = []
    for _ in range(2):
        o = empty(1, np.float64)
        fgx.reduce_sum(ctx, 0, ptr(d_a), n, ptr(o), ptr(s), sb)
        sync()
        outs.append(float(host(o, np.float64)[0]))
    assert outs[0] == outs[1]
    assert abs(outs[0] - a.astype(np.float64).sum()) < 1e-6 * n
    o = empty(1, np.uint64)
    fgx.reduce_sum(ctx, 3, ptr(d_u), n, ptr(o), ptr(s), sb)
    sync()
    assert int(host(o, np.uint64)[0]) == int(u.astype(np.uint64).sum())


@pytest.mark.parametrize("code,dtype,acc", [(0, np.float32, np.float64), (1, np.float64, np.float64),
                                            (2, np.int32, np.int64), (3, np.uint32, np.uint64),
                                            (4, np.int64, np.int64), (5, np.uint64, np.uint64)])
@pytest.mark.parametrize("n", [0, 1, 101, 1_000_003])
def test_reduce_stats(ctx, code, dtype, acc, n):
    """Sum / min / max exact (integers) or to double rounding (floats); the squared
    deviation sum against numpy's two-pass formula; min/max of an empty list are the
    reduction identities."""
    from flamegpu2_amd import fgx
    rng = np.random.default_rng(5 + code)
    if np.issubdtype(dtype, np.floating):
        a = rng.normal(3.0, 40.0, size=n).astype(dtype)
    elif np.issubdtype(dtype, np.signedinteger):
        a = rng.integers(-100000, 100000, size=n).astype(dtype)
    else:
        a = rng.integers(0, 200000, size=n).astype(dtype)
    sb = fgx.scratch_bytes(fgx.FGX_OP_STATS, 0)
    s = scratch(sb)
    d_a = dev(a if n else np.zeros(1, dtype))
    o = empty(4, np.uint64)
    fgx.reduce_stats(ctx, code, ptr(d_a), n, ptr(o), ptr(s), sb)
    sync()
    raw = host(o, np.uint64, 4)
    sm, mn, mx = raw[:3].view(acc)
    sq = raw[3:].view(np.float64)[0]
    if n == 0:
        assert sm == 0 and sq == 0.0
        if acc is np.float64:
            assert mn == np.inf and mx == -np.inf
        else:
            assert mn == np.iinfo(acc).max and mx == np.iinfo(acc).min
        return
    w = a.astype(acc)
    assert mn == w.min() and mx == w.max()
    if acc is np.float64:
        assert abs(sm - w.sum()) <= 1e-12 * np.abs(w).sum()
    else:
        assert int(sm) == int(w.sum())
    mean = float(sm) / n
    d = a.astype(np.float64) - mean
    ref = float((d * d).sum())
    assert abs(sq - ref) <= 1e-10 * max(ref, 1.0)
