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


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, arrs, steps):
    """The serial oracle (g++ -O3, one thread pinned to one core) on `steps` steps of the
    same workload, timed on this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ref = O.Sim(1, params, 7)
    ref.set_agents(0, arrs)
    ref.step(1)   # warm: first sort of an unsorted population
    allowed = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {min(allowed)})
    try:
        t0 = time.perf_counter()
        ref.step(steps)
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, allowed)
    n = len(arrs[0])
    return {"value": n * steps / dt, "unit": "agent-steps/s", "cores": 1, "kind": "port",
            "sample": f"serial C++ oracle (g++ -O3), boids {n} agents, {steps} steps after 1 warm step, one thread "
                      f"pinned to one core, {dt:.1f} s",
            "host_cores": os.cpu_count(), "host_cpu": _cpu_model()}


def ensemble_plans(count=256, seed=2024, n_lo=1024, n_hi=262144, r_lo=1.0, r_hi=3.0):
    """The batched-ensemble workload: circles simulations with log-uniform populations,
    uniform radius, density 1 (width cbrt(N)), one seed each."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        r = float(rng.uniform(r_lo, r_hi))
        out.append((n, r, 1000 + k))
    return out


def bench_batched(args):
    """Batched ensemble vs the same simulations run one after another (plain launches)."""
    from flamegpu2_amd import autotune, workloads
    from flamegpu2_amd.models import Batch, Simulation
    plans = ensemble_plans(args.ensemble)
    steps = args.ensemble_steps
    made = [workloads.circles(n, seed=s, radius=r) for (n, r, s) in plans]
    agents = sum(n for (n, _, _) in plans)
    total = float(agents) * steps
    # batched: all simulations in one batch; the tuner's trial steps are part of the run.
    # The bin-width multiplier is chosen first, on throwaway copies (it changes the
    # summation order, so it cannot change inside the run); its trials count too.
    t0 = time.perf_counter()
    bpr, bw_timings = 1, {}
    if args.autotune:
        def make(k):
            c = Batch("circles_spatial3D", made[0][0])
            c.set_option("bins_per_radius", k)
            for (p, a), (_, _, s) in zip(made, plans):
                c.add(p, s, 8, {"Circle": a})
            return c
        bpr, bw_timings = autotune.choose_bin_width(make)
    b = Batch("circles_spatial3D", made[0][0])
    b.set_option("bins_per_radius", bpr)
    hs = [b.add(p, s, steps, {"Circle": a}) for (p, a), (_, _, s) in zip(made, plans)]
    b.step(1)
    b.sync()
    tuned, timings = None, {}
    done = 1
    if args.autotune and steps - done > autotune.steps_needed(2) + 2:
        tuned, timings = autotune.autotune(b, steps=2)
        done += autotune.steps_needed(2)
    b.step(steps - done)
    for h in hs:
        b.get_agents(h, "Circle")
    wall_b = time.perf_counter() - t0
    print(f"bench: batched {wall_b:.2f} s", file=sys.stderr, flush=True)
    # steady-state step time of the tuned batch (separate short run, same simulations)
    b2 = Batch("circles_spatial3D", made[0][0])
    b2.set_option("bins_per_radius", bpr)
    if tuned:
        b2.set_tuning(*tuned)
    for (p, a), (_, _, s) in zip(made, plans):
        b2.add(p, s, 8, {"Circle": a})
    ms_step_b = b2.timed_steps(4) / 4
    b2.close()
    b.close()
    # plain: each simulation with its own launches, one after another, with the same bin
    # width as the batched run (like for like)
    t0 = time.perf_counter()
    gpu_ms = 0.0
    for (p, a), (_, _, s) in zip(made, plans):
        sim = Simulation("circles_spatial3D", p, seed=s)
        sim.set_option("bins_per_radius", bpr)
        sim.set_agents("Circle", a)
        gpu_ms += sim.timed_steps(steps)
        sim.get_agents("Circle")
        sim.close()
    wall_p = time.perf_counter() - t0
    print(f"bench: plain {wall_p:.2f} s", file=sys.stderr, flush=True)
    return {
        "workload": f"{len(plans)} circles simulations, N log-uniform [1024, 262144] (seeded), r ~ U[1, 3], "
                    f"density 1, {steps} steps each; batched run autotuned over block {{64..1024}} x agents per "
                    f"thread {{1,2,4}} (+ persistent grids of 16/24/32 waves per CU) x bin width {{1.0, 0.5}} x "
                    f"radius; plain runs use the batched run's bin width",
        "agents": agents,
        "batched_agent_steps_per_s": total / wall_b,
        "plain_agent_steps_per_s": total / wall_p,
        "speedup": wall_p / wall_b,
        "plain_stepping_only_agent_steps_per_s": total / (gpu_ms / 1e3),
        "speedup_vs_plain_stepping_only": (gpu_ms / 1e3) / wall_b,
        "batched_wall_s": wall_b,
        "plain_wall_s": wall_p,
        "batched_ms_per_step_steady": ms_step_b,
        "tuning": {"block": tuned[0], "agents_per_thread": tuned[1], "blocks_per_cu": tuned[2],
                   "bin_width_multiplier": 1.0 / bpr} if tuned else None,
        "bin_width_ms_per_step": {f"{1.0 / k:g}": round(v, 4) for k, v in bw_timings.items()},
        "tuning_ms_per_step": {f"{k[0]}x{k[1]}/{k[2]}": round(v, 4) for k, v in timings.items()},
    }


def bench_serving(args):
    from flamegpu2_amd import serving
    reqs = serving.poisson_requests(args.serve_rate, args.serve_seconds, seed=99)
    cap = args.serve_max_agents or None
    max_wait = args.serve_max_wait if args.serve_max_wait > 0 else None
    srv = serving.Server(max_chunk=16, policy=args.serve_policy, max_active_agents=cap, max_wait=max_wait)
    t0 = time.perf_counter()
    res = srv.run(reqs, progress=lambda t: print("bench: " + t, file=sys.stderr, flush=True),
                  stop_after=args.serve_stop_after)
    wall = time.perf_counter() - t0
    out = serving.summarize(res, wall, srv.unfinished)
    out["host_phase_s"] = {k: round(v, 3) for k, v in srv.phase.items()}
    out["workload"] = (f"Poisson {args.serve_rate:g} req/s for {args.serve_seconds:g} s, circles/boids 50/50, "
                       f"N log-uniform [2^10, 2^20], steps U[10, 1000], random seeds; continuous batching")
    if args.serve_seconds < 600:
        out["window"] = f"reduced arrival window ({args.serve_seconds:g} s of the configured 600 s)"
    if args.serve_stop_after is not None:
        out["stopped_after_s"] = args.serve_stop_after
    out["policy"] = args.serve_policy
    out["max_active_agents"] = cap
    out["max_wait_s"] = max_wait
    out["offered_agent_steps_per_s"] = sum(r.n * r.steps for r in reqs) / args.serve_seconds
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "boids_inputdata_traffic.md")
COPY_BYTES = 1 << 30   # calibration copy: 1 GiB read + 1 GiB written


def measured_traffic():
    """HBM bytes per neighbour-iteration launch from the committed PMC summary, or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def traffic_probe(args):
    """The program that the FETCH_SIZE / WRITE_SIZE passes profile: six eager boids
    steps at the bench size (with LDS neighbourhood staging when --staging), then
    three launches of a plain elementwise kernel over
    COPY_BYTES (negation: a copy kernel with known traffic), which calibrates the
    counters."""
    import torch
    from flamegpu2_amd import workloads
    from flamegpu2_amd.models import Simulation
    params, arrs = workloads.boids(args.agents, seed=7)
    sim = Simulation("boids_spatial3D", params, seed=7)
    sim.set_option("graphs", 0)
    sim.set_option("staging", 1 if args.staging else 0)
    sim.set_agents("Boid", arrs)
    sim.step(6)   # the first steps sort a random population; the last ones are the steady state
    sim.sync()
    sim.close()
    print("traffic probe: boids steps done", file=sys.stderr, flush=True)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.neg(src, out=dst)   # a plain elementwise kernel: COPY_BYTES read, COPY_BYTES written
    torch.cuda.synchronize()
    print("traffic probe: copies done", file=sys.stderr, flush=True)


def traffic_report(fetch_dir, write_dir, command):
    """Turn the two counter passes into the committed traffic summary.  Counters are KiB
    per dispatch; each is scaled by what it reads for the calibration copy, whose true
    traffic is COPY_BYTES each way."""
    import csv
    import glob

    def per_dispatch(d, counter):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        vals, names = {}, set()
        for f in files:
            for i, r in enumerate(csv.DictReader(open(f))):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"]
                names.add(name[:120])
                key = "inputdata" if "inputdata" in name else ("copy" if "neg_kernel" in name else None)
                if key:
                    d_id = (f, r.get("Dispatch_Id", i))
                    vals.setdefault(key, {}).setdefault(d_id, 0.0)
                    vals[key][d_id] += float(r["Counter_Value"])
        if set(vals) != {"inputdata", "copy"}:
            raise SystemExit(f"{counter}: kernels found: {sorted(names)}")
        return {k: (sum(v.values()) / len(v), len(v)) for k, v in vals.items()}

    fetch = per_dispatch(fetch_dir, "FETCH_SIZE")
    write = per_dispatch(write_dir, "WRITE_SIZE")
    f_scale = fetch["copy"][0] * 1024 / COPY_BYTES
    w_scale = write["copy"][0] * 1024 / COPY_BYTES
    f_bytes = fetch["inputdata"][0] * 1024 / f_scale
    w_bytes = write["inputdata"][0] * 1024 / w_scale
    with open(TRAFFIC_FILE, "w") as o:
        o.write("# HBM traffic of the neighbour-iteration kernel (boids, 1,048,576 agents)\n\n")
        o.write("Commands (one counter pass each):\n\n")
        o.write("```\n" + command + "\n```\n\n")
        o.write("| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB |\n|---|---|---|---|\n")
        for k in ("inputdata", "copy"):
            o.write(f"| {k} | {fetch[k][1]} | {fetch[k][0]:.1f} | {write[k][0]:.1f} |\n")
        o.write(f"\nCalibration: the copy moves {COPY_BYTES} B each way; FETCH_SIZE reads "
                f"{f_scale:.4f} x and WRITE_SIZE {w_scale:.4f} x the true bytes.\n\n")
        o.write(f"fetch_bytes_per_launch: {f_bytes:.0f}\n\nwrite_bytes_per_launch: {w_bytes:.0f}\n\n")
        o.write(f"bytes_per_launch: {f_bytes + w_bytes:.0f}\n")
    print(open(TRAFFIC_FILE).read())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--agents", type=int, default=1 << 20)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ensemble", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--no-batched", dest="batched", action="store_false")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-policy", choices=("fifo", "shortest"), default="shortest",
                    help="admission order: arrival, or least work (agents x steps) first")
    ap.add_argument("--serve-max-agents", type=int, default=2_000_000,
                    help="cap on the agents of the runs resident at once; 0: HBM budget only")
    ap.add_argument("--serve-max-wait", type=float, default=30.0,
                    help="a run queued longer than this many seconds is admitted first; 0: no aging")
    ap.add_argument("--serve-stop-after", type=float, default=None,
                    help="end serving after this many seconds (unfinished runs reported, latencies as lower bounds)")
    ap.add_argument("--serving-only", action="store_true", help="run only the serving leg and print its summary")
    ap.add_argument("--no-serving", dest="serving", action="store_false")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run the short program profiled by the FETCH_SIZE/WRITE_SIZE passes")
    ap.add_argument("--staging", action="store_true", help="traffic probe: stage neighbourhoods in LDS")
    ap.add_argument("--traffic-report", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "COMMAND"),
                    help="write the traffic summary from two rocprofv3 --pmc output folders")
    args = ap.parse_args()
    if args.traffic_report:
        traffic_report(*args.traffic_report)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (single-GPU build).", file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.serving_only:
        print(json.dumps({"serving": bench_serving(args)}))
        return

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
    traffic = measured_traffic()

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
            # `achieved` counts the algorithmic bytes of the fixed definition (every candidate
            # re-read at 28 B), which the caches serve: the neighbour walk reads each candidate
            # as one packed 16-byte record, mostly from L2, so the figure exceeds the HBM peak.
            # The measured DRAM rate is traffic / launch time; the kernel is limited by the
            # texture-address unit's gathers and by vector instruction issue
            "achieved_basis": "algorithmic bytes per launch (28 B per candidate, the fixed definition) / mean launch time",
            "measured_dram_frac": (traffic / avg_s / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "binding_limit": "gather address processing and vector issue, not DRAM (profiles/boids_1M_kernel_counters.md, "
                             "profiles/boids_1M_wave_cycles.md)",
            "kernel": "agent_function_kernel_w8<inputdata_impl> (neighbour iteration, 8 waves per SIMD)",
            "algorithmic_bytes_per_agent": round(bytes_per_agent, 1),
            "mean_launch_ms": avg_s * 1e3,
        },
        "kernels_ms_per_step": {k: round(v[0] / max(v[1], 1), 4) for k, v in sorted(kt.items())},
    }
    # PCIe-inclusive rate (not `value`): the same steps plus uploading the population
    # from host arrays and reading the final state back (fgxm_set_agents / fgxm_get_agents)
    sim.sync()
    t_dn = time.perf_counter()
    sim.get_agents("Boid")
    t_dn = time.perf_counter() - t_dn
    t_up = time.perf_counter()
    sim.set_agents("Boid", arrs)
    sim.sync()
    t_up = time.perf_counter() - t_up
    out["pcie_inclusive"] = {"upload_s": round(t_up, 5), "download_s": round(t_dn, 5),
                             "agent_steps_per_s": n * args.steps / (ms / 1e3 + t_up + t_dn)}
    sim.close()
    if args.batched:
        print("bench: batched ensemble", file=sys.stderr, flush=True)
        out["batched"] = bench_batched(args)
    if args.serving:
        print("bench: serving", file=sys.stderr, flush=True)
        out["serving"] = bench_serving(args)
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(params, arrs, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
