"""State files on the GPU: the reference's circles input state loaded into the
circles model and stepped against the oracle from the same state; exported states
(XML and JSON) resume bit-exactly, including a model with births, deaths and random
draws (the step counter is restored from the file)."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle_lib as O
from flamegpu2_amd import workloads
from flamegpu2_amd.models import Simulation

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "circles_spatial3D_0.xml")
PARAMS = np.array([10.0, 1.0, 0.05], dtype=np.float32)


def _fixture_arrays():
    root = ET.parse(FIXTURE).getroot()
    recs = [{c.tag: c.text.strip() for c in a} for a in root.findall("xagent")]
    ids = np.array([int(r["_id"]) for r in recs], dtype=np.uint32)
    arrs = [np.array([np.float32(r[v]) for r in recs], dtype=np.float32) for v in ("x", "y", "z", "drift")]
    return ids, arrs


def test_reference_circles_state_loads_and_steps():
    ids, arrs = _fixture_arrays()
    sim = Simulation("circles_spatial3D", PARAMS, seed=1)
    sim.load_state(FIXTURE)
    got, gid = sim.get_agents("Circle")
    assert np.array_equal(gid, ids)
    for g, e in zip(got, arrs):
        assert np.array_equal(g, e)
    ref = O.Sim(0, PARAMS, 1)
    ref.set_agents(0, arrs)
    sim.step(1)
    ref.step(1)
    got, gid = sim.get_agents("Circle")
    exp, eid = ref.get_agents(0, [np.float32] * 4)
    # the oracle numbers IDs 1..n in input order, as the fixture does
    assert np.array_equal(gid, eid)
    for g, e in zip(got, exp):
        np.testing.assert_allclose(g, e, rtol=1e-5, atol=1e-6)


def _state(sim, names):
    out = {}
    for n in names:
        arrs, ids = sim.get_agents(n)
        out[n] = (ids, arrs)
    return out


def _same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k][0], b[k][0]), f"{k}: IDs differ"
        for x, y in zip(a[k][1], b[k][1]):
            assert np.array_equal(x, y), f"{k}: values differ"


@pytest.mark.parametrize("ext", ["xml", "json"])
def test_boids_export_resume_bit_exact(tmp_path, ext):
    params, arrs = workloads.boids(20000, seed=4)
    a = Simulation("boids_spatial3D", params, seed=4)
    a.set_agents("Boid", arrs)
    a.step(3)
    path = tmp_path / f"boids.{ext}"
    a.export_state(path)
    a.step(4)
    b = Simulation("boids_spatial3D", params, seed=4)
    b.load_state(path)
    assert b.step_counter() == 3
    b.step(4)
    _same(_state(a, ["Boid"]), _state(b, ["Boid"]))


@pytest.mark.parametrize("ext", ["xml", "json"])
def test_predprey_export_resume_bit_exact(tmp_path, ext):
    params, prey, pred = workloads.predprey(4000, 1000, seed=3, width=60.0)
    a = Simulation("predator_prey", params, seed=3)
    a.set_agents("Prey", prey)
    a.set_agents("Predator", pred)
    a.step(5)
    path = tmp_path / f"pp.{ext}"
    a.export_state(path, pretty=(ext == "xml"))
    a.step(6)
    b = Simulation("predator_prey", params, seed=3)
    b.load_state(path)
    assert b.step_counter() == 5
    b.step(6)
    _same(_state(a, ["Prey", "Predator"]), _state(b, ["Prey", "Predator"]))
