"""Serving loop: simulation runs arrive over time and are served by continuous batching.

Each request is one simulation run (model, population size, step count, seed).
Requests are admitted while the HBM they need (agent, message, bin-table and
scratch bytes, from the batch's own estimate) fits the budget, and, when a cap
on resident agents is set, while the agents fit the cap.  Policy "fifo" admits
in arrival order; policy "shortest" admits the queued run with the least work
(agents x steps) first, which shortens typical latency under overload.  So that
long runs are not starved, a run that has waited longer than `max_wait` seconds
is admitted before any other, oldest first, and while such a run does not fit
nothing else is admitted (the capacity that frees up is kept for it).  Under
"shortest", a run that does not fit the agent cap does not block the queue: the
next candidate that fits is admitted instead.
An admitted run joins its model's batch at the next step and shares every launch
with the runs already resident, so runs that arrive while others are on the GPU
overlap with them.  The batches of different models run on their own streams.
A run leaves its batch once its steps are done and its final population has been
read back; that read-back is its completion time.

Results do not depend on co-scheduling: a run's final state equals the same run
alone (ordered IDs, Philox keyed by seed / agent ID / step, stable orders).
"""
import collections
import time

import numpy as np

from . import workloads
from .models import Batch


def circles_request(n, seed, radius=2.0):
    params, arrs = workloads.circles(n, seed=seed, radius=radius)
    return params, {"Circle": arrs}


def boids_request(n, seed, radius=0.05, density=32768.0):
    # the 1M-agent density (2^20 agents in a 3.2-wide cube), width a whole number of bins
    bins = max(3, int(round((n / density) ** (1.0 / 3.0) / radius)))
    params, arrs = workloads.boids(n, seed=seed, width=bins * radius, radius=radius)
    return params, {"Boid": arrs}


MODELS = {
    "circles": ("circles_spatial3D", "Circle", circles_request),
    "boids": ("boids_spatial3D", "Boid", boids_request),
}

Request = collections.namedtuple("Request", "id arrival model n steps seed")


def poisson_requests(rate, seconds, seed, n_range=(1 << 10, 1 << 20), steps_range=(10, 1000), models=("circles", "boids")):
    """Poisson arrivals at `rate` per second over `seconds`; model chosen uniformly,
    population log-uniform in n_range, steps uniform in steps_range, random seed."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        out.append(Request(len(out), t, models[int(rng.integers(len(models)))], n,
                           int(rng.integers(steps_range[0], steps_range[1] + 1)), int(rng.integers(1 << 31))))
    return out


class Server:
    POLICIES = ("fifo", "shortest")

    def __init__(self, hbm_budget=160 << 30, max_chunk=16, tuning=None, device=0, policy="fifo",
                 max_active_agents=None, max_wait=None):
        if policy not in self.POLICIES:
            raise ValueError(f"unknown admission policy '{policy}' (one of {', '.join(self.POLICIES)})")
        if max_active_agents is not None and max_active_agents <= 0:
            raise ValueError("max_active_agents must be positive")
        if max_wait is not None and max_wait < 0:
            raise ValueError("max_wait must be non-negative")
        self.budget = hbm_budget
        self.policy = policy
        self.max_active_agents = max_active_agents
        self.max_wait = max_wait
        self.max_chunk = max_chunk
        self.tuning = tuning or {}
        self.device = device
        self.batches = {}

    def _batch(self, model):
        if model not in self.batches:
            lib, agent, make = MODELS[model]
            params, _ = make(1024, 0)
            b = Batch(lib, params, device=self.device)
            if model in self.tuning:
                b.set_tuning(*self.tuning[model])
            self.batches[model] = b
        return self.batches[model]

    def _candidates(self, queue, now):
        """Queued runs in admission order, and how many of the first ones are overdue
        (waited longer than max_wait)."""
        overdue = [] if self.max_wait is None else [q for q in queue if now - q.arrival > self.max_wait]
        overdue.sort(key=lambda q: (q.arrival, q.id))
        late = {q.id for q in overdue}
        rest = [q for q in queue if q.id not in late]
        if self.policy == "fifo":
            rest.sort(key=lambda q: (q.arrival, q.id))
        else:
            rest.sort(key=lambda q: (q.n * q.steps, q.arrival, q.id))
        return overdue + rest, len(overdue)

    def run(self, requests, collect=False, clock=time.perf_counter, progress=None, stop_after=None):
        """Serves `requests` (sorted by arrival, seconds from the start).  Returns one
        record per completed request: dict(id, model, n, steps, arrival, admitted,
        completed, latency[, agents, ids]).  `progress(text)` is called every ~5 s of
        serving.  With `stop_after` (seconds), serving ends at that time even if runs are
        still queued or active; `self.unfinished` then lists them as dict(id, n, steps,
        arrival, admitted or None, steps_done)."""
        self.phase = collections.Counter()
        self.unfinished = []
        last_report = clock()
        reqs = sorted(requests, key=lambda r: r.arrival)
        queue = collections.deque()
        active = {}        # (model, handle) -> [request, bytes, admitted, steps issued]
        used = 0
        agents = 0         # agents of the active runs
        results = []
        made = {}          # request id -> (params, populations) while it waits for HBM
        nxt = 0
        t0 = clock()
        while nxt < len(reqs) or queue or active:
            now = clock() - t0
            if stop_after is not None and now >= stop_after:
                break
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                queue.append(reqs[nxt])
                nxt += 1
            # admission, by policy, while the HBM budget and the agent cap allow; a run
            # that does not fit waits (it is always admitted when nothing is active)
            order, n_overdue = self._candidates(queue, now)
            for pos, r in enumerate(order):
                if active and self.max_active_agents is not None and agents + r.n > self.max_active_agents:
                    if pos < n_overdue or self.policy == "fifo":
                        break      # keep the freed capacity for the overdue / first-come run
                    continue       # shortest: a smaller run further on may fit
                lib, agent, make = MODELS[r.model]
                b = self._batch(r.model)
                tp = clock()
                if r.id not in made:
                    made[r.id] = make(r.n, r.seed)
                params, pops = made[r.id]
                h = b.add(params, r.seed, r.steps, pops)
                need = b.estimate_bytes(h)
                self.phase["admit_s"] += clock() - tp
                if used + need > self.budget and active:
                    b.remove(h)
                    break          # HBM is the hard limit: wait for completions
                del made[r.id]
                used += need
                agents += r.n
                active[(r.model, h)] = [r, need, clock() - t0, 0]
                queue.remove(r)
            if not active:
                if nxt < len(reqs):
                    time.sleep(max(0.0, reqs[nxt].arrival - (clock() - t0)))
                continue
            # one chunk of steps per batch: up to the first completion in it
            for model, b in self.batches.items():
                mine = [v for (m, _), v in active.items() if m == model]
                if not mine:
                    continue
                k = min(min(v[0].steps - v[3] for v in mine), self.max_chunk)
                if k <= 0:
                    continue
                tp = clock()
                b.step(k)
                self.phase["step_call_s"] += clock() - tp
                self.phase["steps"] += k
                for v in mine:
                    v[3] += k
            # completions: read back, then leave the batch
            for model, b in self.batches.items():
                for h in b.finished():
                    key = (model, h)
                    if key not in active:
                        continue
                    tp = clock()
                    r, need, admitted, _ = active.pop(key)
                    rec = {"id": r.id, "model": r.model, "n": r.n, "steps": r.steps, "arrival": r.arrival,
                           "admitted": admitted}
                    if collect:
                        arrays, ids = b.get_agents(h, MODELS[model][1])
                        rec["agents"], rec["ids"] = arrays, ids
                    else:
                        b.sync()
                    rec["completed"] = clock() - t0
                    rec["latency"] = rec["completed"] - r.arrival
                    b.remove(h)
                    used -= need
                    agents -= r.n
                    results.append(rec)
                    self.phase["complete_s"] += clock() - tp
            if progress and clock() - last_report > 5.0:
                last_report = clock()
                progress(f"serving t={clock() - t0:.1f}s arrived={nxt}/{len(reqs)} queued={len(queue)} "
                         f"active={len(active)} done={len(results)} {dict(self.phase)}")
        for v in active.values():
            r = v[0]
            self.unfinished.append({"id": r.id, "n": r.n, "steps": r.steps, "arrival": r.arrival, "admitted": v[2],
                                    "steps_done": v[3]})
        for r in list(queue) + list(reqs[nxt:]):
            self.unfinished.append({"id": r.id, "n": r.n, "steps": r.steps, "arrival": r.arrival, "admitted": None,
                                    "steps_done": 0})
        return results


def summarize(results, wall, unfinished=()):
    """Latency percentiles of the completed runs, the longest wait for admission, and the
    agent-steps served per second of wall time (steps of unfinished runs included).  With
    unfinished runs (a run cut off by stop_after) it also gives the percentiles over every
    arrived request, counting an unfinished one at its wait so far: a lower bound."""
    lat = np.array([r["latency"] for r in results]) if results else np.zeros(1)
    waits = [r["admitted"] - r["arrival"] for r in results]
    waits += [(u["admitted"] if u["admitted"] is not None else wall) - u["arrival"] for u in unfinished
              if u["arrival"] <= wall]
    agent_steps = float(sum(r["n"] * r["steps"] for r in results))
    partial = float(sum(u["n"] * u["steps_done"] for u in unfinished))
    out = {
        "requests": len(results),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "max_wait_s": float(max(waits)) if waits else 0.0,
        "agent_steps": agent_steps + partial,
        "agent_steps_per_s": (agent_steps + partial) / wall if wall > 0 else 0.0,
        "wall_s": wall,
    }
    if unfinished:
        arrived = [u for u in unfinished if u["arrival"] <= wall]
        low = np.concatenate([lat if results else np.zeros(0), np.array([wall - u["arrival"] for u in arrived])])
        out["unfinished"] = len(unfinished)
        out["unfinished_arrived"] = len(arrived)
        out["p50_latency_lower_bound_s"] = float(np.percentile(low, 50)) if len(low) else 0.0
        out["p99_latency_lower_bound_s"] = float(np.percentile(low, 99)) if len(low) else 0.0
    return out
