"""CUDAEnsemble restated from tests/test_cases/simulation/test_cuda_ensemble.cu, through the
model library's ensemble entry (fgxm_ensemble_simulate):

  simulate (:279-322): init functions create 32 agents per run, one agent function counts
  steps, an exit function sums the counters; runs of 1 and 2 steps.  Batchable, so it runs
  as one batch by default and as separate simulations with batched=0; both must agree.
  ErrorOff / ErrorSlow / ErrorFast (:653-748): a step function fails every second run; Off
  reports the count, Slow runs every plan then raises, Fast stops after the first failure.
"""
import numpy as np
import pytest

from flamegpu2_amd.models import ensemble

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("batched", [1, 0])
def test_ensemble_simulate(batched):
    res = ensemble("kat_models", [], seeds=[11, 12], steps=[1, 2], variant="ensemble_simulate",
                   options=f"batched={batched}", sum_var="EXIT_SUM")
    assert res["error"] == "", res["error"]
    assert res["batched"] == batched
    assert res["status"].tolist() == [0, 0]
    assert res["counts"][:, 0].tolist() == [32, 32]
    # expected result: sum over runs of steps x population
    assert res["sums"].tolist() == [32.0, 64.0] and res["sums"].sum() == 96.0


def test_ensemble_simulate_many_runs_batched_equals_separate():
    seeds = list(range(40))
    steps = [1 + (i % 7) for i in range(40)]
    a = ensemble("kat_models", [], seeds=seeds, steps=steps, variant="ensemble_simulate", sum_var="EXIT_SUM")
    b = ensemble("kat_models", [], seeds=seeds, steps=steps, variant="ensemble_simulate", options="batched=0;concurrent=4",
                 sum_var="EXIT_SUM")
    assert a["batched"] == 1 and b["batched"] == 0
    assert a["error"] == "" and b["error"] == ""
    assert np.array_equal(a["sums"], b["sums"])
    assert np.array_equal(a["sums"], 32.0 * np.array(steps))


def _errors(level):
    return ensemble("kat_models", [], seeds=list(range(10)), steps=[1] * 10, variant="ensemble_errors",
                    options=f"error_level={level};concurrent=1")


def test_ensemble_error_off():
    res = _errors(0)
    assert res["error"] == "" and res["batched"] == 0
    assert (res["status"] == 1).sum() == 5 and (res["status"] == 0).sum() == 5   # runs with odd seeds fail


def test_ensemble_error_slow():
    res = _errors(1)
    assert "EnsembleError" in res["error"]
    assert (res["status"] == 1).sum() == 5 and (res["status"] == 2).sum() == 0   # every run ran


def test_ensemble_error_fast():
    res = _errors(2)
    assert "EnsembleError" in res["error"]
    # run 0 succeeds, run 1 fails, nothing else starts
    assert res["status"].tolist()[:2] == [0, 1] and (res["status"] == 2).sum() == 8
