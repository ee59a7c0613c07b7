"""Benchmark: agent-steps/s of spatial-3D boids at 1,048,576 agents on one MI355X.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20]

Prints one JSON line.  The workload is the boids configuration (2^20 agents
uniform in a 3.2-wide cube, |v| = 0.05, interaction radius 0.05 -> 64^3 bins,
wrap-around, seed 7) run through the FLAME GPU 2 compatible C++ API
(flamegpu2_amd/examples/boids_spatial3D.hip).  A step is one full simulation
step: agent spatial sort, message output, message-list build (PBM), neighbour
iteration + movement.  Inputs are resident in HBM before the timed region;
the K timed steps are bracketed by device synchronisation and timed with HIP
events on the simulation stream (whole steps replayed as hipGraphs).

roofline: the dominant kernel is the neighbour-iteration agent function
(inputdata).  Its algorithmic bytes per agent are
    S_in_r + 8*S + C*M + S_w = 28 + 8*27 + C*28 + 24
(own agent read, 27 wrapped cells x 2 PBM bounds, C candidate messages of 28 B,
agent write), C measured exactly from the PBM.  achieved = N * bytes / mean
launch time from HIP events around each launch (eager steps).

cpu_baseline: the serial CPU oracle (oracle/liboracle.so, one thread) on a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "agent-steps/s at 1M boids (+% HBM roofline); batched vs plain speedup; p99 latency"


def wrap_candidates(pbm, grid):
    """Exact mean candidate count of the wrapped 27-cell iterator: sum_b n_b * sum_{27 nbrs} n_b' / N."""
    counts = np.diff(pbm.astype(np.int64)).reshape(grid, grid, grid)   # [z, y, x]
    n = counts.sum()
    if n == 0:
        return 0.0
    nb = np.zeros_like(counts)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                nb += np.roll(counts, shift=(dz, dy, dx), axis=(0, 1, 2))
    return float((counts * nb).sum()) / float(n)


def cpu_baseline(params, arrs, steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ref = O.Sim(1, params, 7)
    ref.set_agents(0, arrs)
    ref.step(1)   # warm: first sort of an unsorted population
    t0 = time.perf_counter()
    ref.step(steps)
    dt = time.perf_counter() - t0
    n = len(arrs[0])
    return {"value": n * steps / dt, "unit": "agent-steps/s", "cores": 1, "kind": "port",
            "sample": f"serial C++ oracle, boids {n} agents, {steps} steps after 1 warm step, one thread, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--agents", type=int, default=1 << 20)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (single-GPU build).", file=sys.stderr)
        sys.exit(2)

    from flamegpu2_amd import workloads
    from flamegpu2_amd.models import Simulation

    n = args.agents
    params, arrs = workloads.boids(n, seed=7)
    sim = Simulation("boids_spatial3D", params, seed=7)
    sim.set_agents("Boid", arrs)
    grid = int(round((float(params[1]) - float(params[0])) / float(params[2])))

    # warm-up (captures the step graphs)
    sim.step(args.warmup)
    sim.sync()
    c0 = wrap_candidates(sim.pbm("location"), grid)
    wall0 = time.perf_counter()
    ms = sim.timed_steps(args.steps)
    sim.sync()
    wall = time.perf_counter() - wall0
    c1 = wrap_candidates(sim.pbm("location"), grid)
    value = n * args.steps / (ms / 1e3)

    # per-kernel durations from HIP events around every launch (eager steps)
    sim.set_option("kernel_timing", 1)
    sim.kernel_times(reset=True)
    ca = wrap_candidates(sim.pbm("location"), grid)
    sim.step(args.timing_steps)
    kt = sim.kernel_times(reset=True)
    cb = wrap_candidates(sim.pbm("location"), grid)
    sim.set_option("kernel_timing", 0)
    fn_ms, fn_cnt = kt["fn:Boid::inputdata"]
    avg_s = fn_ms / fn_cnt / 1e3
    cand = 0.5 * (ca + cb)
    bytes_per_agent = 28 + 8 * 27 + cand * 28 + 24
    achieved = n * bytes_per_agent / avg_s / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "boids_inputdata_traffic.json")
    if os.path.exists(pmc):
        try:
            traffic = json.load(open(pmc)).get("bytes_per_launch")
        except Exception:
            traffic = None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "boids spatial-3D wrapped, 1,048,576 agents, 3.2-wide cube, r=0.05 (64^3 bins), |v|=0.05, seed 7",
            "agents": n,
            "candidates_per_agent_start": round(c0, 2),
            "candidates_per_agent_end": round(c1, 2),
            "graphs": True,
            "wall_seconds_timed": round(wall, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "agent_function_kernel<inputdata_impl> (neighbour iteration)",
            "algorithmic_bytes_per_agent": round(bytes_per_agent, 1),
            "mean_launch_ms": avg_s * 1e3,
        },
        "kernels_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in sorted(kt.items())},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(params, arrs, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
