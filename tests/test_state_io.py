"""State-file reader on the host (no GPU): the reference's circles input state
(examples/cpp/circles_spatial3D/0.xml, kept as tests/golden/circles_spatial3D_0.xml)
parsed against the circles model, checked against an independent parse with
xml.etree; the same content as JSON; malformed files raise."""
import json
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from flamegpu2_amd.fgx import FgxError
from flamegpu2_amd.models import state_summary

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "circles_spatial3D_0.xml")
PARAMS = (10.0, 1.0, 0.05)


def _etree_agents():
    root = ET.parse(FIXTURE).getroot()
    recs = []
    for a in root.findall("xagent"):
        recs.append({c.tag: c.text.strip() for c in a})
    return root, recs


def _expect_sums(recs):
    sums = {}
    for var in ("x", "y", "z", "drift"):
        sums[var] = float(np.float32([np.float32(r[var]) for r in recs]).astype(np.float64).sum())
    sums["_id"] = float(sum(int(r["_id"]) for r in recs))
    return sums


def test_reference_circles_xml_fixture():
    root, recs = _etree_agents()
    assert int(root.find("itno").text) == 0
    s = state_summary("circles_spatial3D", FIXTURE, PARAMS)
    assert s["step"] == 0
    assert s["env"] == {}
    c = s["agents"]["Circle"]["default"]
    assert c["n"] == len(recs) == 1024
    exp = _expect_sums(recs)
    for var, total in exp.items():
        assert c["vars"][var][0] == pytest.approx(total, rel=1e-12, abs=1e-9), var


def test_json_matches_xml(tmp_path):
    _, recs = _etree_agents()
    doc = {"config": {"simulation": {"random_seed": 5, "steps": 3}}, "stats": {"step_count": 7},
           "environment": {"ENV_MAX": 10.0},
           "agents": {"Circle": {"default": [{k: (int(v) if k == "_id" else float(v)) for k, v in r.items() if k != "name"}
                                             for r in recs]}}}
    p = tmp_path / "state.json"
    p.write_text(json.dumps(doc))
    s = state_summary("circles_spatial3D", p, PARAMS)
    assert s["step"] == 7
    assert s["env"]["ENV_MAX"] == [10.0]
    c = s["agents"]["Circle"]["default"]
    assert c["n"] == 1024
    for var, total in _expect_sums(recs).items():
        assert c["vars"][var][0] == pytest.approx(total, rel=1e-12, abs=1e-9), var


@pytest.mark.parametrize("text", ["<states><xagent><name>Circle</name><x>1</x></states>",
                                  "<states><xagent><x>1</x></xagent></states>",
                                  "<states><xagent><name>Nobody</name><x>1</x></xagent></states>",
                                  "<states><xagent><name>Circle</name><nosuchvar>1</nosuchvar></xagent></states>",
                                  "<states><xagent><name>Circle</name><x>one</x></xagent></states>",
                                  "<states><environment><nosuchprop>1</nosuchprop></environment></states>"])
def test_malformed_xml_raises(tmp_path, text):
    p = tmp_path / "bad.xml"
    p.write_text(text)
    with pytest.raises(FgxError):
        state_summary("circles_spatial3D", p, PARAMS)


def test_unknown_extension_raises(tmp_path):
    p = tmp_path / "state.txt"
    p.write_text("{}")
    with pytest.raises(FgxError):
        state_summary("circles_spatial3D", p, PARAMS)
