"""The reference's own known-answer tests, restated through the drop-in C++ API
(tests/models/kat_models.hip) and run on the GPU.  Expected values are computed
on the host exactly as the reference tests compute them."""
import numpy as np
import pytest

from flamegpu2_amd.fgx import FgxError
from flamegpu2_amd.models import Simulation

pytestmark = pytest.mark.gpu


def _sim(variant, params=(), seed=1):
    return Simulation("kat_models", params, seed=seed, variant=variant)


def _expected_counts(P, dims, grid, mask=None):
    """test_spatial_3d.cu:127-171: per bin, messages in the clamped Moore neighbourhood."""
    cells = np.floor(P).astype(np.int64).clip(0, grid - 1)
    occ = np.zeros((grid,) * dims, dtype=np.int64)
    src = cells if mask is None else cells[:, mask]
    np.add.at(occ, tuple(src), 1)
    pad = np.pad(occ, 1)
    nb = np.zeros_like(occ)
    for off in np.ndindex(*([3] * dims)):
        sl = tuple(slice(o, o + grid) for o in off)
        nb += pad[sl]
    return nb[tuple(cells)]


# spatial tests run with the per-wave unit tables (default) and with the per-lane unit walk
@pytest.mark.parametrize("tables", [1, 0])
@pytest.mark.parametrize("variant", ["mandatory3d", "optional3d", "optionalnone3d"])
def test_spatial3d_message_counts(variant, tables):
    rng = np.random.default_rng(11)
    n = 2049
    P = rng.uniform(0, 5, size=(3, n)).astype(np.float32)
    do_output = (rng.uniform(size=n) < 0.8).astype(np.int32)
    s = _sim(variant)
    s.set_option("unit_tables", tables)
    z = np.zeros(n, np.uint32)
    s.set_agents("agent", [P[0], P[1], P[2], do_output, z, z])
    s.step(1)
    (x, y, zz, dout, count, bad), ids = s.get_agents("agent")
    Q = np.stack([x, y, zz])
    if variant == "mandatory3d":
        expect = _expected_counts(Q, 3, 5)
    elif variant == "optional3d":
        expect = _expected_counts(Q, 3, 5, mask=dout == 1)
    else:
        expect = np.zeros(n, np.int64)
    assert np.array_equal(count.astype(np.int64), expect)
    assert np.all(bad == 0)


@pytest.mark.parametrize("tables", [1, 0])
def test_spatial2d_message_counts(tables):
    rng = np.random.default_rng(12)
    n = 2049
    P = rng.uniform(0, 5, size=(2, n)).astype(np.float32)
    s = _sim("mandatory2d")
    s.set_option("unit_tables", tables)
    z = np.zeros(n, np.uint32)
    s.set_agents("agent", [P[0], P[1], z, z])
    s.step(1)
    (x, y, count, bad), _ = s.get_agents("agent")
    assert np.array_equal(count.astype(np.int64), _expected_counts(np.stack([x, y]), 2, 5))
    assert np.all(bad == 0)


@pytest.mark.parametrize("tables", [1, 0])
@pytest.mark.parametrize("offset", [(0.0, 0.0, 0.0), (141.0, -540.0, 200.0), (-1401.5, 5640.3, -2008.8)])
def test_spatial3d_wrapped(offset, tables):
    off = np.array(offset, np.float32)
    i, j, k = np.meshgrid(np.arange(35), np.arange(35), np.arange(35), indexing="ij")
    P = [(i.ravel() * np.float32(2.0)).astype(np.float32) + off[0],
         (j.ravel() * np.float32(2.0)).astype(np.float32) + off[1],
         (k.ravel() * np.float32(2.0)).astype(np.float32) + off[2]]
    n = len(P[0])
    s = _sim("wrapped3d", list(offset))
    s.set_option("unit_tables", tables)
    zf, zu = np.zeros(n, np.float32), np.zeros(n, np.uint32)
    s.set_agents("agent", P + [zf, zf, zf, zu, zu])
    s.step(1)
    (x, y, z, rx, ry, rz, count, bad), _ = s.get_agents("agent")
    assert np.all(rx == 0) and np.all(ry == 0) and np.all(rz == 0)
    assert np.all(bad <= 189)
    assert np.all(count == 27)


@pytest.mark.parametrize("tables", [1, 0])
@pytest.mark.parametrize("offset", [(0.0, 0.0), (141.0, -540.0), (-1401.5, 5640.3)])
def test_spatial2d_wrapped(offset, tables):
    off = np.array(offset, np.float32)
    i, j = np.meshgrid(np.arange(35), np.arange(35), indexing="ij")
    P = [(i.ravel() * np.float32(2.0)).astype(np.float32) + off[0],
         (j.ravel() * np.float32(2.0)).astype(np.float32) + off[1]]
    n = len(P[0])
    s = _sim("wrapped2d", list(offset))
    s.set_option("unit_tables", tables)
    zf, zu = np.zeros(n, np.float32), np.zeros(n, np.uint32)
    s.set_agents("agent", P + [zf, zf, zu, zu])
    s.step(1)
    (x, y, rx, ry, count, bad), _ = s.get_agents("agent")
    assert np.all(rx == 0) and np.all(ry == 0)
    assert np.all(count == 9)


def test_spatial3d_bounds_not_factor_radius():
    h = np.float32(50.1) / np.float32(2)
    pts = np.array([[h, 0.0, h], [h, 18.0, h], [0.0, h, 0.5], [18.0, h, 0.5],
                    [h, 50.0, 50.1], [h, 50.0, np.float32(50.1) - np.float32(10.11)]], dtype=np.float32).T
    s = _sim("bounds3d")
    s.set_agents("agent", [pts[0], pts[1], pts[2], np.zeros(6, np.uint32)])
    s.step(1)
    (_, _, _, count), ids = s.get_agents("agent")
    for c, i in zip(count, ids):
        assert c == (2 if i < 5 else 1)


@pytest.mark.parametrize("period,expect", [(1, [3, 2, 1, 0]), (0, [0, 1, 2, 3])])
def test_auto_spatial_sort(period, expect):
    x = -np.arange(4, dtype=np.float32)
    s = _sim("sort3d", [float(period)])
    s.set_agents("agent", [np.arange(4, dtype=np.int32), x, x, x])
    s.step(1)
    (order, _, _, _), _ = s.get_agents("agent")
    assert order.tolist() == expect


@pytest.mark.parametrize("variant,n_total,n1,n12", [
    ("birth_mandatory", 2048, 1024, 1024),
    ("birth_optional", 1536, 1024, 512),
    ("birth_mandatory_death", 1024, 0, 1024),
    ("birth_optional_death", 1024, 512, 512),
])
def test_device_agent_creation(variant, n_total, n1, n12):
    n = 1024
    s = _sim(variant)
    s.set_agents("agent", [np.arange(n, dtype=np.float32) + 1.0, np.arange(n, dtype=np.uint32)])
    s.step(1)
    (x, idv), ids = s.get_agents("agent")
    assert len(x) == n_total
    val = x - idv.astype(np.float32)
    assert int((val == 1.0).sum()) == n1
    assert int((val == 12.0).sum()) == n12
    # _id: unique; births numbered after the parents, in parent order
    assert len(set(ids.tolist())) == len(ids)
    born = ids[val == 12.0]
    assert np.array_equal(born, np.sort(born)) and born.min() == n + 1


def test_agent_id_multiple_states_unique_ids():
    # DeviceAgentCreationTest.AgentID_MultipleStatesUniqueIDs (test_device_agent_creation.cu:1204-1296):
    # every agent of states a and b gives birth and stores its child's ID (agent_out.getID());
    # the child stores its parent's ID.  IDs must be unique, the in-model copies must match
    # the exported IDs, and parent/child pairings must agree both ways.
    n = 100
    s = _sim("agent_id_states")
    z = np.zeros(n, np.uint32)
    s.set_agents("agent", [z, z], state="a")
    s.set_agents("agent", [z, z], state="b")
    s.step(1)
    pops = [s.get_agents("agent", st) for st in ("a", "b")]
    assert [len(p[1]) for p in pops] == [2 * n, 2 * n]
    ids = np.concatenate([p[1] for p in pops])
    copy = np.concatenate([p[0][0] for p in pops])
    other = np.concatenate([p[0][1] for p in pops])
    assert len(set(ids.tolist())) == len(ids)
    assert np.array_equal(ids, copy)
    pair = {int(i): int(o) for i, o in zip(ids, other) if o != 0}
    assert len(pair) == len(ids)
    for i, o in pair.items():
        assert pair[o] == i


def test_agent_id_birth_into_other_agent_type():
    # AgentID_DeviceBirth_MultipleAgents (test_device_agent_creation.cu:1297-1339): agents A give
    # birth to agents B, whose IDs must come from B's counter (no collision with B's own)
    s = _sim("agent_id_two_types")
    s.set_agents("agentA", [np.zeros(100, np.uint32)])
    s.set_agents("agentB", [np.zeros(200, np.uint32)])
    s.step(1)
    (bo,), bids = s.get_agents("agentB")
    (ao,), aids = s.get_agents("agentA")
    assert len(bids) == 300 and len(set(bids.tolist())) == 300
    # the children carry their parents' IDs, and each parent holds its child's ID
    child_of = dict(zip(aids.tolist(), ao.tolist()))
    for bid, parent in zip(bids.tolist(), bo.tolist()):
        if parent:
            assert child_of[parent] == bid


def test_state_transition():
    n = 1000
    s = _sim("transition")
    s.set_agents("agent", [np.arange(n, dtype=np.uint32)], state="a")
    s.step(1)
    assert s.count("agent", "a") == 0 and s.count("agent", "b") == n
    (x,), _ = s.get_agents("agent", "b")
    assert np.array_equal(x, np.arange(n, dtype=np.uint32) + 1)


# tests/test_cases/runtime/agent/detail/test_agent_state_transition.cu: a transition into an
# empty state swaps the two lists (CUDAFatAgent.cu:139-184); kernel timing labels the
# transitions "transition-swap:" (lists swapped) or "transition:" (agents appended)


def _transitions(s):
    kt = s.kernel_times(reset=True)
    return {k: v[1] for k, v in kt.items() if k.startswith("transition")}


@pytest.mark.parametrize("graphs", [0, 1])
def test_transition_src0_dest10(graphs):
    n = 10
    s = _sim("transition_src0_dest10")
    s.set_option("graphs", graphs)
    y0 = np.tile(np.array([13, 14, 15, 16], np.int32), n)
    s.set_agents("agent", [np.full(n, 12, np.int32), y0], state="Start")
    for step in range(2):
        s.step(1)
        assert s.count("agent", "Start") == 0 and s.count("agent", "End") == n and s.count("agent", "End2") == 0
        (x, y), _ = s.get_agents("agent", "End")
        assert np.all(x == 11) and np.array_equal(y.reshape(n, 4), np.tile([23, 24, 25, 26], (n, 1)))


@pytest.mark.parametrize("graphs", [0, 1])
def test_transition_src10_dest0(graphs):
    n = 10
    s = _sim("transition_src10_dest0")
    s.set_option("graphs", graphs)
    y0 = np.tile(np.array([13, 14, 15, 16], np.int32), n)
    s.set_agents("agent", [np.full(n, 12, np.int32), y0], state="Start")
    s.step(1)
    assert [s.count("agent", st) for st in ("Start", "End", "End2")] == [0, n, 0]
    (x, y), _ = s.get_agents("agent", "End")
    assert np.all(x == 11) and np.array_equal(y.reshape(n, 4), np.tile([23, 24, 25, 26], (n, 1)))
    s.step(1)
    assert [s.count("agent", st) for st in ("Start", "End", "End2")] == [0, 0, n]
    (x, y), ids = s.get_agents("agent", "End2")
    assert np.all(x == 13) and np.array_equal(y.reshape(n, 4), np.tile([3, 4, 5, 6], (n, 1)))
    assert sorted(ids.tolist()) == list(range(1, n + 1))


def test_transition_swaps_only_into_an_empty_state():
    n = 1000
    s = _sim("transition_src10_dest0")
    s.set_option("kernel_timing", 1)
    s.set_agents("agent", [np.arange(n, dtype=np.int32), np.zeros(4 * n, np.int32)], state="Start")
    s.step(1)
    # End2 <- End (nothing to move, End2 empty), then End <- Start (End empty): both swap
    assert _transitions(s) == {"transition-swap:Function2": 1, "transition-swap:Function1": 1}
    # agents placed in Start again, End holds n: Function2 empties End into the empty End2
    # first (swap); Function1 then moves Start into the now-empty End (swap)
    (x, y), ids = s.get_agents("agent", "End")
    s.set_agents("agent", [x, y], ids=ids + n, state="Start")
    s.step(1)
    assert _transitions(s) == {"transition-swap:Function2": 1, "transition-swap:Function1": 1}
    assert [s.count("agent", st) for st in ("Start", "End", "End2")] == [0, n, n]
    # End2 now holds n: the next End -> End2 transition appends
    s.step(1)
    assert _transitions(s) == {"transition:Function2": 1, "transition-swap:Function1": 1}
    assert [s.count("agent", st) for st in ("Start", "End", "End2")] == [0, 0, 2 * n]
    (x, y), ids = s.get_agents("agent", "End2")
    assert np.all(x == 13) and len(set(ids.tolist())) == 2 * n


@pytest.mark.parametrize("graphs", [0, 1])
def test_transition_src10_dest10_conditional(graphs):
    """Src_10_Dest_10: agents leave Start as their counter reaches zero, 10 per step."""
    n, rounds = 10, 3
    s = _sim("transition_src10_dest10")
    s.set_option("graphs", graphs)
    val = (1 + np.arange(rounds * n) % rounds).astype(np.uint32)
    z = np.tile(np.array([13, 14, 15, 16], np.int32), rounds * n)
    s.set_agents("agent", [val, val.copy(), z], state="Start")
    for i in range(1, rounds + 1):
        s.step(1)
        assert s.count("agent", "Start") == (rounds - i) * n and s.count("agent", "End") == i * n
        (_, ys, zs), _ = s.get_agents("agent", "Start")
        assert np.array_equal(zs.reshape(-1, 4), np.tile([23, 24, 25, 26], (len(ys), 1)))
        assert sorted(set(ys.tolist())) == list(range(i + 1, rounds + 1))
        (_, ye, ze), _ = s.get_agents("agent", "End")
        assert np.array_equal(ze.reshape(-1, 4), np.tile([3, 4, 5, 6], (len(ye), 1)))
        assert np.array_equal(np.bincount(ye, minlength=rounds + 1)[1:i + 1], [n] * i)


# tests/test_cases/simulation/test_cuda_simulation_concurrency.cu:76-107,353-390
# (LayerConcurrency): 4 agent types x 512 agents, one slow function each in one layer;
# the step must run at least 1.5x faster with in-layer concurrency than without.  With
# layer fusion the layer is one launch; without it, one launch per function on forked streams.


def _mean_step_ms(s, pops, reps=3):
    total = 0.0
    for r in range(reps + 1):
        for name, arr in pops:
            s.set_agents(name, [arr])
        ms = s.timed_steps(1)
        if r:
            total += ms
    return total / reps


@pytest.mark.parametrize("fusion", [1, 0])
def test_layer_concurrency_speedup(fusion):
    s = _sim("concurrency")
    s.set_option("layer_fusion", fusion)
    pops = [(f"agent_{i}", np.arange(512, dtype=np.float32)) for i in range(4)]
    s.set_option("concurrency", 0)
    seq = _mean_step_ms(s, pops)
    s.set_option("concurrency", 1)
    before = s.stat("fused_layer_launches")
    con = _mean_step_ms(s, pops)
    fused = s.stat("fused_layer_launches") - before
    print(f"fusion {fusion}: sequential {seq:.3f} ms, concurrent {con:.3f} ms, speedup {seq / con:.2f}, "
          f"fused launches {fused}")
    assert (fused > 0) == bool(fusion)
    assert seq / con >= 1.5
    for i in range(4):
        (v,), _ = s.get_agents(f"agent_{i}")
        assert v[0] == 0 and np.all(np.isinf(v[1:]))


def test_two_radii_sorts_graphs_and_eager():
    """One agent list sorted, every step, for two spatial inputs of different radii: the
    two sorts share the list's scratch with different bin counts.  Graph replays (several
    steps, with eager reads between) must equal eager steps and a brute-force count."""
    rng = np.random.default_rng(23)
    n = 3000
    P = rng.uniform(0, 10, size=(3, n)).astype(np.float32)
    d = P[:, :, None] - P[:, None, :]
    d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
    exp_near = (d2 <= np.float32(1.0) * np.float32(1.0)).sum(axis=1)
    exp_far = (d2 <= np.float32(2.5) * np.float32(2.5)).sum(axis=1)
    runs = {}
    for graphs in (1, 0):
        s = _sim("two_radii")
        s.set_option("graphs", graphs)
        z = np.zeros(n, np.uint32)
        s.set_agents("agent", [P[0], P[1], P[2], z, z])
        out = []
        for _ in range(4):
            s.step(1)
            (x, y, zz, near, far), ids = s.get_agents("agent")
            order = np.argsort(ids)
            out.append((x[order], near[order], far[order]))
            assert np.array_equal(near[order], exp_near) and np.array_equal(far[order], exp_far)
        runs[graphs] = out
    for a, b in zip(runs[1], runs[0]):
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


def test_bruteforce_sum():
    n = 2000
    rng = np.random.default_rng(4)
    x = rng.integers(1, 10, size=n).astype(np.int32)
    s = _sim("bruteforce")
    z = np.zeros(n, np.int32)
    s.set_agents("agent", [x, z, z])
    s.step(1)
    (_, ssum, _), _ = s.get_agents("agent")
    assert np.all(ssum == int(x.sum()))


def test_agent_random_distinct_and_in_range():
    n = 4096
    s = _sim("random", seed=1234)
    z = np.zeros(n, np.float32)
    s.set_agents("agent", [z, z, np.zeros(n, np.int32), z])
    s.step(1)
    (a, b, c, d), ids = s.get_agents("agent")
    assert np.all((a >= 0) & (a < 1)) and np.all((b >= 0) & (b < 1))
    assert len(np.unique(a)) > 0.99 * n and not np.array_equal(a, b)
    assert np.all((c >= 1) & (c <= 6))
    assert abs(float(d.mean())) < 0.1 and abs(float(d.std()) - 1.0) < 0.1
    # counter-based: same seed -> same draws; different seed -> different draws
    s2 = _sim("random", seed=1234)
    s2.set_agents("agent", [z, z, np.zeros(n, np.int32), z])
    s2.step(1)
    (a2, _, _, _), _ = s2.get_agents("agent")
    assert np.array_equal(a, a2)
    s3 = _sim("random", seed=99)
    s3.set_agents("agent", [z, z, np.zeros(n, np.int32), z])
    s3.step(1)
    (a3, _, _, _), _ = s3.get_agents("agent")
    assert not np.array_equal(a, a3)


def test_device_error_death_not_enabled():
    s = _sim("death_disabled")
    s.set_agents("agent", [np.zeros(10, np.float32)])
    with pytest.raises(FgxError, match="DeviceError"):
        s.step(1)


def test_device_error_wrap_out_of_bounds():
    s = _sim("wrap_out_of_bounds")
    x = np.array([5.0, 11.0], np.float32)
    s.set_agents("agent", [x, x, x, np.zeros(2, np.uint32)])
    with pytest.raises(FgxError, match="DeviceError"):
        s.step(1)


# ---------------------------------------------------------------- function conditions
# tests/test_cases/runtime/test_agent_function_conditions.cu


def test_condition_split_agents():
    n = 1000
    s = _sim("cond_split")
    x = (np.arange(2 * n) % 2).astype(np.int32)
    y = np.tile(np.array([13, 14, 15, 16], np.int32), 2 * n)
    s.set_agents("agent", [x, y], state="Start")
    s.step(1)
    assert s.count("agent", "Start") == 0
    assert s.count("agent", "End") == n and s.count("agent", "End2") == n
    (x2, y2), ids2 = s.get_agents("agent", "End")
    (x3, y3), ids3 = s.get_agents("agent", "End2")
    assert np.all(x2 == 2) and np.array_equal(y2.reshape(-1, 4), np.tile([3, 4, 5, 6], (n, 1)))
    assert np.all(x3 == -1) and np.array_equal(y3.reshape(-1, 4), np.tile([23, 24, 25, 26], (n, 1)))
    # agents keep their input order within each part
    assert np.array_equal(ids2, np.arange(2, 2 * n + 1, 2)) and np.array_equal(ids3, np.arange(1, 2 * n, 2))


def test_condition_all_disabled():
    n = 1000
    s = _sim("cond_all_disabled")
    s.set_agents("agent", [np.zeros(n, np.int32), np.zeros(4 * n, np.int32)])
    s.step(3)
    (x, y), ids = s.get_agents("agent")
    assert len(x) == n and np.all(x == -3)
    assert np.array_equal(y.reshape(-1, 4), np.tile([23, 24, 25, 26], (n, 1)))
    assert np.array_equal(ids, np.arange(1, n + 1))


@pytest.mark.parametrize("graphs", [True, False])
def test_condition_with_death_births_and_messages(graphs):
    n = 3000
    x = np.arange(n, dtype=np.int32)
    s = _sim("cond_mixed")
    s.set_option("graphs", graphs)
    s.set_agents("agent", [x, np.zeros(n, np.int32)])
    s.step(1)
    (gx, gsum), gid = s.get_agents("agent")
    ids = np.arange(1, n + 1, dtype=np.uint32)
    enabled = x % 3 != 0
    disabled_x, disabled_id = x[~enabled], ids[~enabled]
    ex, eid = x[enabled], ids[enabled]
    surv = ex % 2 == 1
    born_x = ex[ex % 5 == 0] + 1000
    exp_x = np.concatenate([disabled_x, ex[surv], born_x])
    exp_id = np.concatenate([disabled_id, eid[surv], n + 1 + np.arange(len(born_x), dtype=np.uint32)])
    assert np.array_equal(gx, exp_x)
    assert np.array_equal(gid, exp_id)
    assert np.all(gsum == int(ex.sum()))
