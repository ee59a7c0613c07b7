"""Host agent reductions on the GPU (HostAgentAPI::min / max / sum / meanStandardDeviation /
count, include/flamegpu/runtime/agent/HostAgentAPI.cuh), called from a step function of
tests/models/kat_host.hip and read back through the exit log.

Known answers: tests/test_cases/runtime/agent/host_reduction/test_mean_standarddeviation.cu
(0..100 -> standard deviation 29.15476); otherwise numpy on the same seeded values:
integer sums, minima and maxima exact, float sums (accumulated in double) to 1e-12."""
import json

import numpy as np
import pytest

from flamegpu2_amd.models import Simulation

pytestmark = pytest.mark.gpu

TAGS = {"f": np.float32, "d": np.float64, "i": np.int32, "u": np.uint32, "l": np.int64, "ul": np.uint64}


def _run(tmp_path, cols):
    s = Simulation("kat_host", (), seed=1, variant="reductions")
    s.set_agents("agent", [cols[t] for t in TAGS])
    for t in TAGS:
        for r in ("min_", "max_", "sum_", "mean_", "sd_"):
            s.log("env", name=r + t, exit=True)
    s.log("env", name="count", exit=True)
    s.simulate(1)
    s.export_log(tmp_path / "l.json", steps=False, exit=True)
    return json.loads((tmp_path / "l.json").read_text())["exit"]["env"]


def test_reference_mean_standard_deviation(tmp_path):
    n = 101
    cols = {t: np.arange(n).astype(dt) for t, dt in TAGS.items()}
    env = _run(tmp_path, cols)
    assert env["count"] == n
    for t in TAGS:
        assert env["mean_" + t] == 50.0
        assert np.float32(env["sd_" + t]) == pytest.approx(np.float32(29.15476), rel=4 * 2.0 ** -23)
        assert env["min_" + t] == 0 and env["max_" + t] == 100 and env["sum_" + t] == 5050


@pytest.mark.parametrize("n", [1, 777, 300_001])
def test_reductions_match_numpy(tmp_path, n):
    rng = np.random.default_rng(n)
    cols = {"f": rng.normal(0, 50, n).astype(np.float32), "d": rng.normal(5, 1e3, n),
            "i": rng.integers(-10**6, 10**6, n).astype(np.int32), "u": rng.integers(0, 2 * 10**6, n).astype(np.uint32),
            "l": rng.integers(-10**12, 10**12, n).astype(np.int64), "ul": rng.integers(0, 10**12, n).astype(np.uint64)}
    env = _run(tmp_path, cols)
    for t, a in cols.items():
        w = a.astype(np.float64)
        assert env["min_" + t] == float(a.min()) and env["max_" + t] == float(a.max())
        if a.dtype.kind in "iu":
            assert int(env["sum_" + t]) == int(a.astype(object).sum()) or env["sum_" + t] == float(a.astype(object).sum())
        else:
            assert env["sum_" + t] == pytest.approx(w.sum(), rel=1e-12, abs=1e-9)
        assert env["mean_" + t] == pytest.approx(w.mean(), rel=1e-12, abs=1e-9)
        assert env["sd_" + t] == pytest.approx(w.std(), rel=1e-9, abs=1e-9)
