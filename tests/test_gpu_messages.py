"""Bucket and array messages on the GPU (tests/models/kat_messages.hip).

Bucket: the reference's BucketMessageTest cases (tests/test_cases/runtime/messaging/
test_bucket.cu:99-330,612-690) with their own expected counts and sums.
Array: every agent reads its own element through at() and walks a neighbourhood; the
expected element sequence is the reference iterators' visiting order
(MessageArrayDevice.cuh:765-818; MessageArray2DDevice.cuh:1366-1540;
MessageArray3DDevice.cuh:1536-1571): x offset outermost, then y, then z, origin
skipped, clamped to the array or wrapped modulo its dimensions, von Neumann keeping
|dx|+|dy|+|dz| <= r.  The reference's Moore tests (test_array_2d.cu:199-316: 8 and 24
wrapped neighbours) are the r = 1, 2 wrapped cases of the same check."""
import itertools

import numpy as np
import pytest

from flamegpu2_amd.fgx import FgxError
from flamegpu2_amd.models import Simulation

pytestmark = pytest.mark.gpu

AGENT_COUNT = 1024


def _bucket_sim(variant, do_output=None):
    s = Simulation("kat_messages", [AGENT_COUNT], seed=1, variant=variant)
    ids = np.arange(AGENT_COUNT, dtype=np.int32)
    dout = np.ones(AGENT_COUNT, np.int32) if do_output is None else do_output.astype(np.int32)
    z = np.zeros(AGENT_COUNT, np.uint32)
    s.set_agents("agent", [ids, dout, z, z, z])
    s.step(1)
    (rid, _, c1, c2, sm), _ = s.get_agents("agent")
    order = np.argsort(rid)
    return rid[order], c1[order], c2[order], sm[order]


def _bucket_tables(outputs):
    count = np.zeros(AGENT_COUNT // 2, np.int64)
    total = np.zeros(AGENT_COUNT // 2, np.int64)
    for i in range(AGENT_COUNT):
        if outputs[i]:
            count[i // 2] += 1
            total[i // 2] += i
    return count, total


@pytest.mark.parametrize("variant", ["bucket_mandatory", "bucket_optional", "bucket_optional_none"])
def test_bucket(variant):
    rng = np.random.default_rng(3)
    outputs = {"bucket_mandatory": np.ones(AGENT_COUNT, bool),
               "bucket_optional": rng.uniform(size=AGENT_COUNT) > 0.3,
               "bucket_optional_none": np.zeros(AGENT_COUNT, bool)}[variant]
    rid, c1, c2, sm = _bucket_sim(variant, outputs)
    count, total = _bucket_tables(outputs)
    key = np.maximum(rid - 1, 0) // 2
    assert np.array_equal(c1, count[key])
    assert np.array_equal(c2, count[key])
    assert np.array_equal(sm.astype(np.int64), total[key])


def test_bucket_range():
    rid, c1, c2, sm = _bucket_sim("bucket_range")
    count, total = _bucket_tables(np.ones(AGENT_COUNT, bool))
    b = (rid // 8) * 4
    exp_c = sum(count[b + j] for j in range(4))
    exp_s = sum(total[b + j] for j in range(4))
    assert np.array_equal(c1, exp_c)
    assert np.array_equal(c2, count[rid // 2])
    assert np.array_equal(sm.astype(np.int64), exp_s)


def _neighbourhood(shape, loc, r, mode):
    """Elements visited by the reference iterator, in order."""
    D = len(shape)
    wrap = mode in (1, 3)
    vn = mode in (2, 3)
    ranges = []
    for a in range(D):
        if wrap:
            ranges.append(range(-r, r + 1))
        else:
            lo = -loc[a] if loc[a] - r < 0 else -r
            hi = shape[a] - 1 - loc[a] if loc[a] + r >= shape[a] else r
            ranges.append(range(lo, hi + 1))
    out = []
    for d in itertools.product(*ranges):   # x outermost, z innermost
        if all(v == 0 for v in d):
            continue
        if vn and sum(abs(v) for v in d) > r:
            continue
        c = [(loc[a] + d[a]) % shape[a] if wrap else loc[a] + d[a] for a in range(D)]
        e = c[0] + (c[1] * shape[0] if D > 1 else 0) + (c[2] * shape[0] * shape[1] if D > 2 else 0)
        out.append(e)
    return out


def _array_case(shape, radius, mode, optional, seed=5):
    D = len(shape)
    dims = list(shape) + [1] * (3 - D)
    n = int(np.prod(shape))
    s = Simulation("kat_messages", [D, dims[0], dims[1], dims[2], radius, mode, optional], seed=1, variant="array")
    rng = np.random.default_rng(seed)
    write = rng.permutation(n).astype(np.uint32)
    index = np.arange(n, dtype=np.uint32)
    z = np.zeros(n, np.uint32)
    s.set_agents("agent", [index, write, z, z, z])
    s.step(1)
    (idx, _, read, hsh, cnt), _ = s.get_agents("agent")
    order = np.argsort(idx)
    read, hsh, cnt = read[order], hsh[order], cnt[order]
    value = np.array([e * 3 + 1 if (not optional or e % 2 == 0) else 0 for e in range(n)], dtype=np.uint64)
    for i in range(n):
        loc = [i % dims[0], (i // dims[0]) % dims[1], i // (dims[0] * dims[1])][:D]
        assert read[i] == value[i], f"at({loc}) read {read[i]}, expected {value[i]}"
        seq = _neighbourhood(shape, loc, radius, mode)
        h = 0
        for e in seq:
            h = (h * 31 + int(value[e])) & 0xFFFFFFFF
        assert cnt[i] == len(seq), f"element {loc}: {cnt[i]} neighbours, expected {len(seq)}"
        assert hsh[i] == h, f"element {loc}: neighbourhood order or values differ"
    return cnt


@pytest.mark.parametrize("radius,mode", [(1, 0), (2, 0), (1, 1), (3, 1)])
def test_array_1d(radius, mode):
    _array_case((128,), radius, mode, 0)


@pytest.mark.parametrize("radius,mode", [(1, 0), (2, 0), (1, 1), (2, 1), (1, 2), (2, 2), (1, 3), (2, 3)])
def test_array_2d(radius, mode):
    cnt = _array_case((12, 13), radius, mode, 0)
    if mode == 1:
        # test_array_2d.cu:249-316 (Moore1W / Moore2W): 8 and 24 wrapped neighbours
        assert np.all(cnt == (2 * radius + 1) ** 2 - 1)


@pytest.mark.parametrize("radius,mode", [(1, 0), (2, 0), (1, 1), (2, 1), (1, 2), (2, 3)])
def test_array_3d(radius, mode):
    _array_case((5, 6, 7), radius, mode, 0)


@pytest.mark.parametrize("D", [1, 2, 3])
def test_array_optional_output(D):
    shape = {1: (64,), 2: (9, 7), 3: (5, 4, 3)}[D]
    _array_case(shape, 1, 0, 1)


def test_array_duplicate_output_raises():
    """test_array_2d.cu:318-353 (DuplicateOutputException): ArrayMessageWriteConflict."""
    n = 12 * 13
    s = Simulation("kat_messages", [2, 12, 13, 1, 1, 0, 0], seed=1, variant="array")
    write = np.arange(n, dtype=np.uint32)
    write[13] = 0
    z = np.zeros(n, np.uint32)
    s.set_agents("agent", [np.arange(n, dtype=np.uint32), write, z, z, z])
    with pytest.raises(FgxError, match="ArrayMessageWriteConflict"):
        s.step(1)


def test_array_index_out_of_bounds_raises():
    n = 64
    s = Simulation("kat_messages", [1, 64, 1, 1, 1, 0, 0], seed=1, variant="array")
    write = np.arange(n, dtype=np.uint32)
    write[5] = 64
    z = np.zeros(n, np.uint32)
    s.set_agents("agent", [np.arange(n, dtype=np.uint32), write, z, z, z])
    with pytest.raises(FgxError, match="DeviceError"):
        s.step(1)


# ---------------------------------------------------------------- append / truncate, persistence
# tests/test_cases/runtime/messaging/test_append_truncate.cu:72-310 and
# test_brute_force.cu:517-600, through tests/models/kat_messages.hip.
def _kat(variant):
    return Simulation("kat_messages", [], seed=1, variant=variant)


def test_truncate():
    n = 1024
    s = _kat("truncate")
    s.set_agents("Agent", [np.zeros(n, np.uint32)])
    for _ in range(2):
        s.step(1)
        (count,), _ = s.get_agents("Agent")
        assert np.all(count == n)


def test_append_keep_data():
    # two output layers append to one list: every agent reads both sets, each step
    n = 1024
    s = _kat("append_keep")
    z = np.zeros(n, np.uint32)
    s.set_agents("Agent", [z, z])
    for _ in range(2):
        s.step(1)
        (c0, c1), _ = s.get_agents("Agent")
        assert np.all(c0 == n) and np.all(c1 == n)


def test_optional_truncate():
    n = 1024
    do_out = (np.random.default_rng(5).uniform(size=n) < 0.7).astype(np.uint32)
    k = int(do_out.sum())
    s = _kat("optional_truncate")
    s.set_agents("Agent", [np.zeros(n, np.uint32), do_out])
    s.step(1)
    (count, _), _ = s.get_agents("Agent")
    assert np.all(count == k)
    s.step(1)   # the outputting agents flipped do_out: the others output now
    (count, _), _ = s.get_agents("Agent")
    assert np.all(count == n - k)


def test_optional_append_keep_data():
    n = 1024
    do_out = (np.random.default_rng(6).uniform(size=n) < 0.7).astype(np.uint32)
    k = int(do_out.sum())
    s = _kat("optional_append")
    z = np.zeros(n, np.uint32)
    s.set_agents("Agent", [z, z, do_out])
    s.step(1)
    (c0, c1, _), _ = s.get_agents("Agent")
    assert np.all(c0 == k) and np.all(c1 == n - k)
    s.step(1)
    (c0, c1, _), _ = s.get_agents("Agent")
    assert np.all(c0 == n - k) and np.all(c1 == k)


def test_append_keep_data_resize():
    # the second output list is twice the first, forcing the list to grow before the append
    # (the reference's issue #725): the first output must survive the resize
    n = 1024
    s = _kat("append_resize")
    s.set_agents("a", [], n=n)
    s.set_agents("b", [], n=2 * n)
    s.set_agents("c", [np.zeros(1, np.uint32)] * 3)
    s.step(1)
    (c0, c1, c2), _ = s.get_agents("c")
    assert (int(c0[0]), int(c1[0]), int(c2[0])) == (0, n, 2 * n)


@pytest.mark.parametrize("persistent", [False, True])
def test_message_persistence(persistent):
    # output on even steps only: a non-persistent list is empty on odd steps, a persistent
    # one still holds the even step's messages
    n = 128
    s = _kat("persistence_on" if persistent else "persistence_off")
    z = np.zeros(n, np.uint32)
    s.set_agents("agent", [np.arange(n, dtype=np.int32), z, z])
    s.step(1)
    (_, count, total), _ = s.get_agents("agent")
    assert np.all(count == n) and np.all(total == n * (n - 1) // 2)
    s.step(1)
    (_, count, total), _ = s.get_agents("agent")
    if persistent:
        assert np.all(count == n) and np.all(total == n * (n - 1) // 2)
    else:
        assert np.all(count == 0) and np.all(total == 0)
