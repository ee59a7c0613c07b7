"""Admission logic of the serving loop (flamegpu2_amd/serving.py) against a stand-in
batch, on the CPU: policy order, the resident-agent cap, the HBM budget."""
import pytest

from flamegpu2_amd import serving


class StubBatch:
    """Steps every member together, like the batched simulation; no device work."""

    def __init__(self, bytes_per_agent=1):
        self.members = {}
        self.next = 0
        self.bytes_per_agent = bytes_per_agent
        self.max_resident = 0

    def add(self, params, seed, steps, pops):
        h = self.next
        self.next += 1
        self.members[h] = [steps, 0, pops["n"]]
        return h

    def estimate_bytes(self, h):
        return self.members[h][2] * self.bytes_per_agent

    def remove(self, h):
        del self.members[h]

    def step(self, k):
        # the server adds a run to measure it and removes it again when it does not fit,
        # so residency is what a step sees
        self.max_resident = max(self.max_resident, sum(m[2] for m in self.members.values()))
        for m in self.members.values():
            m[1] += k

    def finished(self):
        return [h for h, m in self.members.items() if m[1] >= m[0]]

    def sync(self):
        pass


@pytest.fixture
def stub(monkeypatch):
    b = StubBatch()
    monkeypatch.setattr(serving, "MODELS", {"m": ("lib", "A", lambda n, seed: (None, {"n": n}))})
    monkeypatch.setattr(serving.Server, "_batch", lambda self, model: self.batches.setdefault(model, b))
    return b


def _reqs(spec):
    return [serving.Request(i, 0.0, "m", n, steps, i) for i, (n, steps) in enumerate(spec)]


def _order(res):
    return [r["id"] for r in sorted(res, key=lambda r: (r["completed"], r["id"]))]


def test_fifo_one_at_a_time_keeps_arrival_order(stub):
    reqs = _reqs([(50, 9), (10, 2), (30, 5)])
    res = serving.Server(max_chunk=4, max_active_agents=50).run(reqs)
    assert _order(res) == [0, 1, 2]
    assert stub.max_resident <= 50


def test_shortest_first_orders_by_work(stub):
    reqs = _reqs([(50, 9), (10, 2), (30, 5), (10, 40)])   # work 450, 20, 150, 400
    res = serving.Server(max_chunk=4, policy="shortest", max_active_agents=50).run(reqs)
    assert _order(res) == [1, 2, 3, 0]
    assert stub.max_resident <= 50


def test_cap_admits_an_oversized_run_alone(stub):
    reqs = _reqs([(80, 3), (10, 3)])
    res = serving.Server(max_chunk=4, max_active_agents=50).run(reqs)
    assert sorted(r["id"] for r in res) == [0, 1]
    assert stub.max_resident == 80   # never with another run


def test_no_cap_admits_everything_at_once(stub):
    reqs = _reqs([(50, 9), (10, 2), (30, 5)])
    res = serving.Server(max_chunk=4, policy="shortest").run(reqs)
    assert len(res) == 3
    assert stub.max_resident == 90


def test_hbm_budget_still_applies(stub):
    reqs = _reqs([(40, 3), (40, 3), (40, 3)])
    serving.Server(hbm_budget=100, max_chunk=4).run(reqs)
    assert stub.max_resident == 80


def test_bad_arguments():
    with pytest.raises(ValueError):
        serving.Server(policy="lifo")
    with pytest.raises(ValueError):
        serving.Server(max_active_agents=0)


def test_shortest_does_not_block_on_a_run_that_does_not_fit(stub):
    # cap 50: R0 (30 agents) is admitted; R1 (30) does not fit beside it, R2 (15) does and is
    # admitted at once instead of waiting behind R1
    reqs = _reqs([(30, 2), (30, 3), (15, 10)])   # work 60, 90, 150
    res = serving.Server(max_chunk=1, policy="shortest", max_active_agents=50).run(reqs)
    adm = {r["id"]: r["admitted"] for r in res}
    assert adm[2] < adm[1]
    assert stub.max_resident <= 50


def test_overdue_runs_go_first_and_hold_capacity(stub):
    # with max_wait 0 every queued run is overdue at once: admission falls back to arrival
    # order, and a first run that does not fit holds the queue (no smaller run overtakes)
    reqs = _reqs([(30, 2), (30, 3), (15, 10)])
    res = serving.Server(max_chunk=1, policy="shortest", max_active_agents=50, max_wait=0.0).run(reqs)
    adm = {r["id"]: r["admitted"] for r in res}
    assert adm[0] <= adm[1] <= adm[2]
    # without aging, shortest-first admits by work
    reqs = _reqs([(50, 9), (10, 2), (30, 5), (10, 40)])
    res = serving.Server(max_chunk=4, policy="shortest", max_active_agents=50, max_wait=None).run(reqs)
    assert _order(res) == [1, 2, 3, 0]


def test_stop_after_reports_unfinished_runs(stub):
    reqs = _reqs([(30, 2), (30, 3)])
    srv = serving.Server(max_chunk=1)
    res = srv.run(reqs, stop_after=0.0)
    assert res == [] and sorted(u["id"] for u in srv.unfinished) == [0, 1]
    s = serving.summarize(res, 1.0, srv.unfinished)
    assert s["requests"] == 0 and s["unfinished"] == 2 and s["max_wait_s"] >= 1.0
    assert s["p99_latency_lower_bound_s"] >= 1.0


def test_summary_reports_longest_wait():
    res = [{"id": 0, "n": 10, "steps": 2, "arrival": 0.0, "admitted": 0.5, "completed": 1.0, "latency": 1.0},
           {"id": 1, "n": 10, "steps": 2, "arrival": 0.2, "admitted": 3.0, "completed": 4.0, "latency": 3.8}]
    s = serving.summarize(res, 4.0)
    assert abs(s["max_wait_s"] - 2.8) < 1e-9 and s["agent_steps"] == 40.0


def test_bad_max_wait():
    with pytest.raises(ValueError):
        serving.Server(max_wait=-1.0)
