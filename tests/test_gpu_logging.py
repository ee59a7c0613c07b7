"""Step and exit logs: the reference's LoggingTest cases (tests/test_cases/io/test_logging.cu)
restated through the drop-in C++ API (kat_models "logging" variants), read back through
the JSON log export.  Expected values are the reference test's own."""
import json

import numpy as np
import pytest

from flamegpu2_amd.models import Simulation

pytestmark = pytest.mark.gpu

N = 101


def _logging_sim(seed=1):
    s = Simulation("kat_models", (), seed=seed, variant="logging")
    f = np.arange(N, dtype=np.float32)
    i = np.arange(1, N + 1, dtype=np.int32)
    u = np.arange(2, N + 2, dtype=np.uint32)
    s.set_agents("Agent1", [f, i, u])
    return s


def _configure(s, exit):
    # test_logging.cu:40-46,66-78: count, every reduction of every variable, all properties, timing
    s.log("count", agent="Agent1", exit=exit)
    for var in ("float_var", "int_var", "uint_var"):
        for what in ("min", "max", "mean", "standard_deviation", "sum"):
            s.log(what, agent="Agent1", name=var, exit=exit)
    for p in ("float_prop", "int_prop", "uint_prop", "float_prop_array", "int_prop_array", "uint_prop_array"):
        s.log("env", name=p, exit=exit)
    s.log("timing", value=1, exit=exit)


def _check_frame(fr, k):
    """test_logging.cu:117-157 for a frame at step k."""
    assert fr["step_index"] == k
    (a,) = fr["agents"]["Agent1"]["default"]
    assert a["count"] == N
    v = a["variables"]
    assert v["float_var"]["max"] == 100.0 + k and v["int_var"]["max"] == 101 + k and v["uint_var"]["max"] == 102 + k
    assert v["float_var"]["min"] == 0.0 + k and v["int_var"]["min"] == 1 + k and v["uint_var"]["min"] == 2 + k
    assert v["float_var"]["mean"] == 50.0 + k and v["int_var"]["mean"] == 51.0 + k and v["uint_var"]["mean"] == 52.0 + k
    for var in ("float_var", "int_var", "uint_var"):
        # EXPECT_FLOAT_EQ(29.15476f, ...): within 4 float ulps
        assert np.float32(v[var]["standard_deviation"]) == pytest.approx(np.float32(29.15476), rel=4 * 2.0 ** -23)
    assert v["float_var"]["sum"] == (50.0 + k) * N
    assert v["int_var"]["sum"] == (51 + k) * N
    assert v["uint_var"]["sum"] == (52 + k) * N
    e = fr["env"]
    assert e["float_prop"] == 1.0 + k and e["int_prop"] == 1 + k and e["uint_prop"] == 1 + k
    assert e["float_prop_array"] == [1.0 + k, 2.0 + k]
    assert e["int_prop_array"] == [2 + k, 3 + k, 4 + k]
    assert e["uint_prop_array"] == [3 + k, 4 + k, 5 + k, 6 + k]


def test_step_log_with_step(tmp_path):
    """LoggingTest.CUDASimulationStep: frequency 2 over 10 calls of step(); no frame 0."""
    s = _logging_sim()
    _configure(s, exit=False)
    s.log("frequency", value=2)
    for i in range(1, 11):
        s.step(1)
        s.export_log(tmp_path / "l.json", steps=True, exit=False, step_time=True)
        steps = json.loads((tmp_path / "l.json").read_text())["steps"]
        assert len(steps) == i // 2
    for j, fr in enumerate(steps):
        _check_frame(fr, 2 * (j + 1))
        assert fr["step_time"] >= 0.0


def test_step_and_exit_log_with_simulate(tmp_path):
    """LoggingTest.CUDASimulationSimulate: frame 0 (initial state) + 5 frames, exit log at step 10."""
    s = _logging_sim()
    _configure(s, exit=False)
    _configure(s, exit=True)
    s.log("frequency", value=2)
    s.simulate(10)
    s.export_log(tmp_path / "l.json", steps=True, exit=True, step_time=True, exit_time=True)
    log = json.loads((tmp_path / "l.json").read_text())
    assert len(log["steps"]) == 6
    for j, fr in enumerate(log["steps"]):
        _check_frame(fr, 2 * j)
    _check_frame(log["exit"], 10)
    for k in ("rtc_time", "init_time", "exit_time", "total_time"):
        assert log["exit"][k] >= 0.0
    assert log["config"]["random_seed"] == 1


@pytest.mark.parametrize("graphs", [0, 1])
def test_step_log_graphs_invariant(tmp_path, graphs):
    """Logged frames are the same whether steps replay a captured graph or run eagerly."""
    s = _logging_sim()
    s.set_option("graphs", graphs)
    _configure(s, exit=False)
    s.simulate(6)
    s.export_log(tmp_path / "l.json", steps=True, exit=False)
    steps = json.loads((tmp_path / "l.json").read_text())["steps"]
    assert [f["step_index"] for f in steps] == list(range(7))
    for k, fr in enumerate(steps):
        _check_frame(fr, k)


def test_empty_mean(tmp_path):
    """LoggingTest.EmptyMean: the mean of an empty population logs 0, not NaN."""
    s = Simulation("kat_models", (), seed=1, variant="logging_empty")
    s.log("count", agent="Agent1")
    s.log("mean", agent="Agent1", name="float_var")
    s.simulate(5)
    s.export_log(tmp_path / "l.json", steps=True, exit=False)
    steps = json.loads((tmp_path / "l.json").read_text())["steps"]
    assert len(steps) == 6
    for fr in steps:
        (a,) = fr["agents"]["Agent1"]["default"]
        assert a["count"] == 0
        assert a["variables"]["float_var"]["mean"] == 0.0


def test_log_config_errors():
    """src/flamegpu/simulation/LoggingConfig.cu / AgentLoggingConfig.cu validation."""
    from flamegpu2_amd.fgx import FgxError
    s = Simulation("kat_models", (), seed=1, variant="logging")
    with pytest.raises(FgxError):
        s.log("env", name="no_such_property")
    with pytest.raises(FgxError):
        s.log("mean", agent="Agent1", name="no_such_var")
    with pytest.raises(FgxError):
        s.log("count", agent="NoAgent")
    with pytest.raises(FgxError):
        s.log("count", agent="Agent1", state="no_such_state")
    s.log("env", name="int_prop")
    with pytest.raises(FgxError):
        s.log("env", name="int_prop")
    s.log("sum", agent="Agent1", name="int_var")
    with pytest.raises(FgxError):
        s.log("sum", agent="Agent1", name="int_var")
