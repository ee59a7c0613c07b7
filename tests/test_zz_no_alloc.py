"""fgx_spatial_build_noalloc: the message-list build working in the library's own per-GPU
space, for callers that cannot allocate.  Through the C-ABI, it must give exactly what
fgx_spatial_build gives with a caller scratch buffer (and what the oracle gives), on
consecutive calls that reuse the shared space with a smaller input the second time, and
refuse sizes beyond its limits."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import dev, empty, host, ptr, sync, scratch

pytestmark = pytest.mark.gpu

MAX_MESSAGES = 2097152
MAX_BINS = 2097152


@pytest.fixture(scope="module")
def ctx():
    from flamegpu2_amd import fgx
    c = fgx.Context(0)
    yield c
    c.close()


def _positions(dims, n, width, seed, clump=0.0):
    rng = np.random.default_rng(seed)
    P = rng.uniform(0, width, size=(dims, n)).astype(np.float32)
    if clump:
        k = int(clump * n)
        P[:, rng.permutation(n)[:k]] = np.float32(width / 2)   # one large bin
    return P


def _both_builds(ctx, dims, P, width, radius, use_dn):
    """Runs the no-allocation build and the scratch-buffer build on the same inputs and
    returns both results (bin table, ids, positions) and the oracle's."""
    from flamegpu2_amd import fgx
    n = P.shape[1]
    ids = (np.arange(n, dtype=np.uint32) * 5 + 1).astype(np.uint32)
    mn, mx = [0.0] * dims, [float(width)] * dims
    m = fgx.spatial_meta(dims, mn, mx, radius)
    d_in = [dev(P[a]) for a in range(dims)] + [dev(ids)]
    z = ptr(d_in[2]) if dims == 3 else None
    d_n = dev(np.array([n], np.uint32)) if use_dn else None
    cap = n + 1000 if use_dn else n          # with a device count the capacity may exceed it
    out = {}
    for which in ("noalloc", "scratch"):
        d_out = [empty(cap, np.float32) for _ in range(dims)] + [empty(cap, np.uint32)]
        pbm = empty(m.bins + 1, np.uint32)
        args = (ctx, m, cap, ptr(d_in[0]), ptr(d_in[1]), z, [ptr(t) for t in d_in], [ptr(t) for t in d_out],
                [4] * (dims + 1), ptr(pbm))
        if which == "noalloc":
            fgx.spatial_build_noalloc(*args, d_n=ptr(d_n) if use_dn else None)
        else:
            sb = fgx.scratch_bytes(fgx.FGX_OP_SPATIAL_BUILD, cap, m.bins)
            s = scratch(sb)
            fgx.spatial_build(*args, ptr(s), sb, d_n=ptr(d_n) if use_dn else None)
        sync()
        out[which] = (host(pbm, np.uint32, m.bins + 1), host(d_out[dims], np.uint32, n),
                      [host(d_out[a], np.float32, n) for a in range(dims)])
    mo = O.spatial_meta(dims, mn, mx, radius)
    pbm_o, _, outs_o = O.spatial_build(mo, P[0], P[1], P[2] if dims == 3 else None, extra=[ids])
    return out, (pbm_o, outs_o[dims], outs_o[:dims])


def _check(out, ref):
    (pa, ia, xa), (pb, ib, xb) = out["noalloc"], out["scratch"]
    assert np.array_equal(pa, pb) and np.array_equal(ia, ib)
    for a, b in zip(xa, xb):
        assert np.array_equal(a, b)
    pbm_o, ids_o, pos_o = ref
    assert np.array_equal(pa, pbm_o) and np.array_equal(ia, ids_o)
    for a, b in zip(xa, pos_o):
        assert np.array_equal(a, b)


def test_noalloc_build_matches_scratch_build_twice(ctx):
    # first call: 1 M messages on the boids grid (64^3 bins), with a device-resident count;
    # second call: a smaller 2-D list whose histogram region overlaps what the first left
    P1 = _positions(3, 1 << 20, 3.2, 7)
    print("noalloc build: 1048576 messages, 64^3 bins, device count", flush=True)
    out, ref = _both_builds(ctx, 3, P1, 3.2, 0.05, use_dn=True)
    _check(out, ref)
    P2 = _positions(2, 50_000, 100.0, 8, clump=0.05)
    print("noalloc build: 50000 messages, 2-D 20^2 bins, one bin of 2500", flush=True)
    out, ref = _both_builds(ctx, 2, P2, 100.0, 5.0, use_dn=False)
    _check(out, ref)


def test_noalloc_build_at_the_limits(ctx):
    # the largest accepted list, with a clumped bin (the big-bin ordering uses the space too)
    P = _positions(3, MAX_MESSAGES, 6.4, 9, clump=0.01)
    print(f"noalloc build: {MAX_MESSAGES} messages, 128^3 bins", flush=True)
    out, ref = _both_builds(ctx, 3, P, 6.4, 0.05, use_dn=False)
    _check(out, ref)


def test_noalloc_build_refuses_sizes_beyond_its_limits(ctx):
    from flamegpu2_amd import fgx
    # refused before any device work: nothing is read or written
    pbm = empty(1, np.uint32)
    m = fgx.spatial_meta(3, [0, 0, 0], [3.2, 3.2, 3.2], 0.05)
    with pytest.raises(fgx.FgxError, match="exceed the limits"):
        fgx.spatial_build_noalloc(ctx, m, MAX_MESSAGES + 1, 0, 0, 0, [0], [0], [4], ptr(pbm))
    big = fgx.spatial_meta(3, [0, 0, 0], [1.3, 1.3, 1.3], 0.01)   # 130^3 bins > the bin limit
    assert big.bins > MAX_BINS
    with pytest.raises(fgx.FgxError, match="exceed the limits"):
        fgx.spatial_build_noalloc(ctx, big, 10, 0, 0, 0, [0], [0], [4], ptr(pbm))
