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


def _run_pair(model, oracle_model, params, pops, agent_names, steps, seed, graphs=True, staging=0):
    sim = Simulation(model, params, seed=seed)
    sim.set_option("graphs", graphs)
    sim.set_option("staging", staging)
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


@pytest.mark.parametrize("n,steps,graphs,staging", [(32768, 5, True, 0), (32768, 5, False, 0), (1 << 20, 1, True, 0),
                                                    (32768, 5, True, 1), (1 << 20, 1, True, 1)])
def test_boids_bit_exact(n, steps, graphs, staging):
    params, arrs = workloads.boids(n, seed=7)
    sim, ref = _run_pair("boids_spatial3D", 1, params, [arrs], ["Boid"], steps, 7, graphs=graphs, staging=staging)
    assert np.array_equal(sim.pbm("location"), ref.pbm(0))
    _compare(sim, ref, 0, "Boid", exact=True)


@pytest.mark.parametrize("staging", [0, 1])
def test_predprey_small_bit_exact_traces(staging):
    params, prey, pred = workloads.predprey(40000, 10000, seed=3, width=200.0)
    sim = Simulation("predator_prey", params, seed=3)
    sim.set_option("staging", staging)
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
