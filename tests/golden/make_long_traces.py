"""Writes tests/golden/long_traces.json: long-run traces of the serial oracle at the
configured lengths, for the GPU long-run parity tests.

    python tests/golden/make_long_traces.py            # every trace (boids 1M x 500 steps
                                                       # takes about 10 minutes)
    python tests/golden/make_long_traces.py predprey   # one group

Traces (all from the oracle in oracle/, run in this container):
  circles   16,384 agents, seed 12, 100 steps: mean x, y, z and mean drift at steps
            1, 10, 50, 100.
  boids     1,048,576 agents, seed 7, 500 steps: mean speed |f|, polarisation |sum f| / sum |f|,
            mean x, y, z and the mean candidate count of the step, at steps 100, 250, 500.
  predprey  the full configuration (1,000,000 prey + 250,000 predators in [0, 1000)^2,
            seed 3) for 1,000 steps, and a long-lived variant at the same prey density
            with 100 prey per predator (10,000 prey + 100 predators in [0, 100)^2) whose
            populations oscillate for the whole 1,000 steps.  Per step: both population
            counts and an 8-byte BLAKE2b digest of each population's ID sequence in list
            order.

The ID digest is over the little-endian u32 IDs in list order, so it pins the order
of survivors and births as well as the set.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
from flamegpu2_amd import workloads  # noqa: E402

OUT = os.path.join(HERE, "long_traces.json")


def id_digest(ids):
    return hashlib.blake2b(np.ascontiguousarray(ids, dtype="<u4").tobytes(), digest_size=8).hexdigest()


def circles_trace():
    params, arrs = workloads.circles(16384, seed=12)
    s = O.Sim(0, params, 12)
    s.set_agents(0, arrs)
    marks, out, done = [1, 10, 50, 100], {}, 0
    for m in marks:
        s.step(m - done)
        done = m
        a, _ = s.get_agents(0, [np.float32] * 4)
        out[str(m)] = {"mean_x": float(a[0].astype(np.float64).mean()), "mean_y": float(a[1].astype(np.float64).mean()),
                       "mean_z": float(a[2].astype(np.float64).mean()),
                       "mean_drift": float(a[3].astype(np.float64).mean())}
    return {"workload": "circles spatial-3D, 16384 agents, seed 12", "agents": 16384, "seed": 12, "steps": out}


def boids_trace():
    params, arrs = workloads.boids(1 << 20, seed=7)
    s = O.Sim(1, params, 7)
    s.set_agents(0, arrs)
    marks, out, done = [100, 250, 500], {}, 0
    t0 = time.time()
    for m in marks:
        s.step(m - done)
        done = m
        a, _ = s.get_agents(0, [np.float32] * 6)
        f = [x.astype(np.float64) for x in a[3:6]]
        speed = np.sqrt(f[0] ** 2 + f[1] ** 2 + f[2] ** 2)
        # polarisation |sum v| / sum |v|: 0 for random headings, 1 for one common heading
        pol = float(np.sqrt(f[0].sum() ** 2 + f[1].sum() ** 2 + f[2].sum() ** 2) / speed.sum())
        out[str(m)] = {"mean_speed": float(speed.mean()), "polarisation": pol,
                       "mean_x": float(a[0].astype(np.float64).mean()),
                       "mean_y": float(a[1].astype(np.float64).mean()), "mean_z": float(a[2].astype(np.float64).mean()),
                       "candidates_per_agent": float(s.last_candidates(0))}
        print(f"boids step {m}: {out[str(m)]} ({time.time() - t0:.0f} s)", flush=True)
    return {"workload": "boids spatial-3D wrapped, 1048576 agents, seed 7", "agents": 1 << 20, "seed": 7,
            "steps": out}


def predprey_trace(n_prey, n_pred, width, steps=1000, seed=3):
    params, prey, pred = workloads.predprey(n_prey, n_pred, seed=seed, width=width)
    s = O.Sim(2, params, seed)
    s.set_agents(0, prey)
    s.set_agents(1, pred)
    counts, digests = [], []
    for k in range(steps):
        s.step(1)
        _, ip = s.get_agents(0, [np.float32] * 4)
        _, iq = s.get_agents(1, [np.float32] * 4 + [np.uint32])
        counts.append([int(len(ip)), int(len(iq))])
        digests.append([id_digest(ip), id_digest(iq)])
    return {"workload": f"predator-prey spatial-2D, {n_prey} prey + {n_pred} predators in [0, {width:g})^2, seed {seed}",
            "prey": n_prey, "predators": n_pred, "width": width, "seed": seed, "steps": steps,
            "counts": counts, "id_digests": digests}


def main(groups):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    data["digest"] = "blake2b, 8 bytes, over the little-endian u32 IDs in list order"
    if "circles" in groups:
        data["circles"] = circles_trace()
    if "predprey" in groups:
        data["predprey_full"] = predprey_trace(1_000_000, 250_000, 1000.0)
        data["predprey_sustained"] = predprey_trace(10_000, 100, 100.0)
    if "boids" in groups:
        data["boids"] = boids_trace()
    with open(OUT, "w") as f:
        json.dump(data, f, separators=(",", ":"))
        f.write("\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main(sys.argv[1:] or ["circles", "predprey", "boids"])
