"""Long-run parity at the configured lengths, against oracle traces committed in
tests/golden/long_traces.json (written by tests/golden/make_long_traces.py):

  predator-prey, full configuration (1,000,000 prey + 250,000 predators), 1,000 steps, and
  a long-lived variant (10,000 prey + 100 predators in [0, 100)^2, populations oscillating
  for all 1,000 steps): both population counts and the ID sequence of both populations,
  bit-exact at every step;
  boids, 1,048,576 agents, 500 steps: mean speed, polarisation, mean position and mean
  candidates per agent within 1 % at steps 100, 250 and 500;
  circles, 16,384 agents, 100 steps: mean position and mean drift within 1 %.

1 % is the bar for long runs, where neighbour summation order may legitimately differ; the
ID traces are bit-exact because compaction keeps order and the RNG is keyed by agent ID
and step.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from flamegpu2_amd import workloads
from flamegpu2_amd.models import Simulation

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "long_traces.json")))
REL = 0.01


def _digest(ids):
    return hashlib.blake2b(np.ascontiguousarray(ids, dtype="<u4").tobytes(), digest_size=8).hexdigest()


def _ids(sim, agent):
    return sim.get_agents(agent)[1]


@pytest.mark.parametrize("trace", ["predprey_full", "predprey_sustained"])
def test_predprey_population_and_id_traces(trace):
    g = GOLD[trace]
    params, prey, pred = workloads.predprey(g["prey"], g["predators"], seed=g["seed"], width=g["width"])
    sim = Simulation("predator_prey", params, seed=g["seed"])
    sim.set_agents("Prey", prey)
    sim.set_agents("Predator", pred)
    for k in range(g["steps"]):
        sim.step(1)
        counts = [sim.count("Prey"), sim.count("Predator")]
        assert counts == g["counts"][k], f"{trace} step {k + 1}: populations {counts} vs oracle {g['counts'][k]}"
        if counts[0] or counts[1] or k < 3:
            d = [_digest(_ids(sim, "Prey")), _digest(_ids(sim, "Predator"))]
            assert d == g["id_digests"][k], f"{trace} step {k + 1}: ID sequence differs from the oracle"
    if trace == "predprey_sustained":
        c = np.array(g["counts"])
        assert c[:, 0].max() > 5 * c[:, 0].min() and c[-1].min() > 0   # the trace oscillates and survives


def _wrap_candidates(pbm, grid):
    counts = np.diff(pbm.astype(np.int64)).reshape(grid, grid, grid)
    n = counts.sum()
    nb = np.zeros_like(counts)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nb += np.roll(counts, shift=(dz, dy, dx), axis=(0, 1, 2))
    return float((counts * nb).sum()) / float(n)


def _close(got, exp, what):
    assert abs(got - exp) <= REL * abs(exp), f"{what}: {got} vs oracle {exp} (bar {REL:.0%})"


def test_boids_500_steps_statistics():
    if "boids" not in GOLD:
        pytest.skip("no boids trace in tests/golden/long_traces.json (make_long_traces.py boids)")
    g = GOLD["boids"]
    params, arrs = workloads.boids(g["agents"], seed=g["seed"])
    sim = Simulation("boids_spatial3D", params, seed=g["seed"])
    sim.set_agents("Boid", arrs)
    done = 0
    for mark in sorted(g["steps"], key=int):
        sim.step(int(mark) - done)
        done = int(mark)
        e = g["steps"][mark]
        a, ids = sim.get_agents("Boid")
        assert len(ids) == g["agents"]
        f = [x.astype(np.float64) for x in a[3:6]]
        speed = np.sqrt(f[0] ** 2 + f[1] ** 2 + f[2] ** 2)
        pol = float(np.sqrt(f[0].sum() ** 2 + f[1].sum() ** 2 + f[2].sum() ** 2) / speed.sum())
        _close(float(speed.mean()), e["mean_speed"], f"step {mark} mean speed")
        _close(pol, e["polarisation"], f"step {mark} polarisation")
        for k, ax in enumerate("xyz"):
            _close(float(a[k].astype(np.float64).mean()), e[f"mean_{ax}"], f"step {mark} mean {ax}")
        _close(_wrap_candidates(sim.pbm("location"), 64), e["candidates_per_agent"], f"step {mark} candidates")
        print(f"boids step {mark}: speed {speed.mean():.6f} polarisation {pol:.4f} "
              f"candidates {_wrap_candidates(sim.pbm('location'), 64):.2f}")


def test_circles_100_steps_statistics():
    g = GOLD["circles"]
    params, arrs = workloads.circles(g["agents"], seed=g["seed"])
    sim = Simulation("circles_spatial3D", params, seed=g["seed"])
    sim.set_agents("Circle", arrs)
    done = 0
    for mark in sorted(g["steps"], key=int):
        sim.step(int(mark) - done)
        done = int(mark)
        e = g["steps"][mark]
        a, ids = sim.get_agents("Circle")
        assert len(ids) == g["agents"]
        for k, ax in enumerate("xyz"):
            _close(float(a[k].astype(np.float64).mean()), e[f"mean_{ax}"], f"step {mark} mean {ax}")
        _close(float(a[3].astype(np.float64).mean()), e["mean_drift"], f"step {mark} mean drift")
