"""Whole-model parity on the GPU: the C++ API models (flamegpu2_amd/examples) run
through the model C ABI and are compared with the serial oracle on the same
seeded inputs.

Bars: agent order / IDs / populations / PBM bit-exact; FP32 state bit-exact
where the model uses only correctly rounded operations (boids, predator-prey)
and within 1e-5 relative after one step where it calls sinf (circles)."""
import numpy as np
import pytest

import oracle_lib as O
from flamegpu2_amd import workloads
from flamegpu2_amd.models import Simulation

pytestmark = pytest.mark.gpu


def _run_pair(model, oracle_model, params, pops, agent_names, steps, seed, graphs=True, staging=0, tables=1):
    sim = Simulation(model, params, seed=seed)
    sim.set_option("graphs", graphs)
    sim.set_option("staging", staging)
    sim.set_option("unit_tables", tables)
    ref = O.Sim(oracle_model, params, seed)
    for t, (name, arrs) in enumerate(zip(agent_names, pops)):
        sim.set_agents(name, arrs)
        ref.set_agents(t, arrs)
    sim.step(steps)
    ref.step(steps)
    return sim, ref


def _compare(sim, ref, t, name, exact, rtol=1e-5, atol=1e-6):
    got, gid = sim.get_agents(name)
    dts = [a.dtype for a in got]
    exp, eid = ref.get_agents(t, dts)
    assert len(gid) == len(eid), f"{name}: population {len(gid)} vs oracle {len(eid)}"
    assert np.array_equal(gid, eid), f"{name}: ID sequence differs from the oracle"
    for k, (g, e) in enumerate(zip(got, exp)):
        if exact or g.dtype.kind in "iu":
            if not np.array_equal(g, e):
                bad = np.nonzero(g != e)[0]
                raise AssertionError(f"{name} var {k}: {len(bad)} mismatches, first at {bad[0]}: {g[bad[0]]} vs {e[bad[0]]}")
        else:
            np.testing.assert_allclose(g, e, rtol=rtol, atol=atol)
    return got, gid


def test_circles_config_one_step():
    params, arrs = workloads.circles(16384, seed=12)
    sim, ref = _run_pair("circles_spatial3D", 0, params, [arrs], ["Circle"], 1, 12)
    assert np.array_equal(sim.pbm("location"), ref.pbm(0))
    _compare(sim, ref, 0, "Circle", exact=False)


def test_circles_bruteforce_one_step():
    params, arrs = workloads.circles(2048, seed=5)
    sim, ref = _run_pair("circles_bruteforce", 3, params, [arrs], ["Circle"], 1, 5)
    _compare(sim, ref, 0, "Circle", exact=False)


def test_circles_long_run_statistics():
    params, arrs = workloads.circles(16384, seed=12)
    sim, ref = _run_pair("circles_spatial3D", 0, params, [arrs], ["Circle"], 100, 12)
    got, _ = sim.get_agents("Circle")
    exp, _ = ref.get_agents(0, [np.float32] * 4)
    assert len(got[0]) == 16384
    # FP32 differences from sinf ulps may reorder agents after many steps; compare statistics.
    for k in range(3):
        assert abs(float(got[k].mean()) - float(exp[k].mean())) < 0.01 * float(params[0])
    assert abs(float(got[3].mean()) - float(exp[3].mean())) <= 0.01 * max(float(exp[3].mean()), 1e-6) + 1e-6


# staging: the LDS-staged neighbourhood; tables: per-wave unit tables (0 = per-lane unit walk)
@pytest.mark.parametrize("n,steps,graphs,staging,tables", [(32768, 5, True, 0, 1), (32768, 5, False, 0, 1),
                                                           (1 << 20, 1, True, 0, 1), (32768, 5, True, 1, 1),
                                                           (1 << 20, 1, True, 1, 1), (32768, 5, True, 0, 0),
                                                           (1 << 20, 1, True, 0, 0)])
def test_boids_bit_exact(n, steps, graphs, staging, tables):
    params, arrs = workloads.boids(n, seed=7)
    sim, ref = _run_pair("boids_spatial3D", 1, params, [arrs], ["Boid"], steps, 7, graphs=graphs, staging=staging,
                         tables=tables)
    assert np.array_equal(sim.pbm("location"), ref.pbm(0))
    _compare(sim, ref, 0, "Boid", exact=True)


# fusion: the two functions of each layer as one launch (default) or one launch each
@pytest.mark.parametrize("staging,tables,fusion", [(0, 1, 1), (1, 1, 1), (0, 0, 1), (0, 1, 0)])
def test_predprey_small_bit_exact_traces(staging, tables, fusion):
    params, prey, pred = workloads.predprey(40000, 10000, seed=3, width=200.0)
    sim = Simulation("predator_prey", params, seed=3)
    sim.set_option("staging", staging)
    sim.set_option("unit_tables", tables)
    sim.set_option("layer_fusion", fusion)
    ref = O.Sim(2, params, 3)
    sim.set_agents("Prey", prey)
    sim.set_agents("Predator", pred)
    ref.set_agents(0, prey)
    ref.set_agents(1, pred)
    for step in range(30):
        sim.step(1)
        ref.step(1)
        assert sim.count("Prey") == ref.count(0), f"step {step}: prey population"
        assert sim.count("Predator") == ref.count(1), f"step {step}: predator population"
    _compare(sim, ref, 0, "Prey", exact=True)
    _compare(sim, ref, 1, "Predator", exact=True)
    # layers of two functions run as one launch with fusion on (counted when enqueued)
    assert (sim.stat("fused_layer_launches") > 0) == bool(fusion)


def test_predprey_full_config_three_steps():
    params, prey, pred = workloads.predprey(seed=3)
    sim = Simulation("predator_prey", params, seed=3)
    ref = O.Sim(2, params, 3)
    sim.set_agents("Prey", prey)
    sim.set_agents("Predator", pred)
    ref.set_agents(0, prey)
    ref.set_agents(1, pred)
    for step in range(3):
        sim.step(1)
        ref.step(1)
        assert (sim.count("Prey"), sim.count("Predator")) == (ref.count(0), ref.count(1)), f"step {step}"
    _compare(sim, ref, 0, "Prey", exact=True)
    _compare(sim, ref, 1, "Predator", exact=True)


def test_graph_and_eager_agree():
    params, arrs = workloads.boids(20000, seed=9)
    a = Simulation("boids_spatial3D", params, seed=9)
    b = Simulation("boids_spatial3D", params, seed=9)
    a.set_option("graphs", True)
    b.set_option("graphs", False)
    a.set_agents("Boid", arrs)
    b.set_agents("Boid", arrs)
    a.step(7)
    b.step(7)
    ga, ia = a.get_agents("Boid")
    gb, ib = b.get_agents("Boid")
    assert np.array_equal(ia, ib)
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y)
    assert a.step_counter() == b.step_counter() == 7


# ---------------------------------------------------------------- half-width bins
# Bin-width multiplier 1/2 (option "bins_per_radius" = 2; no counterpart in the
# reference): the oracle runs the same finer grid and visit order (orc_sim_set_bin_mult).
def _run_half(model, oracle_model, params, arrs, name, steps, seed):
    sim = Simulation(model, params, seed=seed)
    sim.set_option("bins_per_radius", 2)
    ref = O.Sim(oracle_model, params, seed, bin_mult=0.5)
    sim.set_agents(name, arrs)
    ref.set_agents(0, arrs)
    sim.step(steps)
    ref.step(steps)
    return sim, ref


def test_half_width_bins_circles():
    params, arrs = workloads.circles(16384, seed=12)
    sim, ref = _run_half("circles_spatial3D", 0, params, arrs, "Circle", 1, 12)
    pbm = sim.pbm("location")
    assert len(pbm) == 26 ** 3 + 1   # ceil(25.398 / 1.0) bins per axis
    assert np.array_equal(pbm, ref.pbm(0))
    _compare(sim, ref, 0, "Circle", exact=False)


@pytest.mark.parametrize("n,steps", [(32768, 4), (1 << 20, 1)])
def test_half_width_bins_boids_bit_exact(n, steps):
    params, arrs = workloads.boids(n, seed=7)
    sim, ref = _run_half("boids_spatial3D", 1, params, arrs, "Boid", steps, 7)
    assert len(sim.pbm("location")) == 128 ** 3 + 1
    assert np.array_equal(sim.pbm("location"), ref.pbm(0))
    _compare(sim, ref, 0, "Boid", exact=True)


def test_half_width_bins_agree_with_full_bins():
    # same in-radius neighbours, different visit order: FP32 agreement after one step
    params, arrs = workloads.boids(32768, seed=8)
    a = Simulation("boids_spatial3D", params, seed=8)
    b = Simulation("boids_spatial3D", params, seed=8)
    b.set_option("bins_per_radius", 2)
    a.set_agents("Boid", arrs)
    b.set_agents("Boid", arrs)
    a.step(1)
    b.step(1)
    (ga, ia), (gb, ib) = a.get_agents("Boid"), b.get_agents("Boid")
    oa, ob = np.argsort(ia), np.argsort(ib)
    assert np.array_equal(ia[oa], ib[ob])
    for x, y in zip(ga, gb):
        np.testing.assert_allclose(x[oa], y[ob], rtol=1e-5, atol=1e-6)


def test_half_width_bins_only_before_first_step():
    from flamegpu2_amd.fgx import FgxError
    params, arrs = workloads.boids(1000, seed=8)
    s = Simulation("boids_spatial3D", params, seed=8)
    s.set_agents("Boid", arrs)
    s.step(1)
    with pytest.raises(FgxError, match="before the first step"):
        s.set_option("bins_per_radius", 2)
    with pytest.raises(FgxError, match="1 or 2"):
        s.set_option("bins_per_radius", 3)
