"""GPU: the autotuner's parameters change speed only. The step launch's
workgroup size leaves every particle, count and footprint bit-identical;
re-binning particles by met cell leaves every particle bit-identical and the
footprint within float32 summation-order rounding (a cell's values are summed
in a different order). The tuner picks one of its candidates, caches it per
workload shape, and a second call returns the cached choice."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd.autotune import tune  # noqa: E402
from glide_amd.runner import BatchEngine  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402

STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "alive")


@pytest.fixture(scope="module")
def setup():
    cfg = dict(workload_config("batched-cpu-6r"), convection=False)
    return cfg, SyntheticMet(cfg["spec"])


def _run(cfg, met, **kw):
    eng = BatchEngine(cfg, device="cuda", met=met, compact=True, **kw)
    eng.run()
    torch.cuda.synchronize()
    return eng


def test_tuning_parameters_leave_results_unchanged(setup):
    cfg, met = setup
    ref = _run(cfg, met)
    S = ref.state_numpy()
    for kw in (dict(step_block=64), dict(step_block=512), dict(step_block=128, reorder_every=16)):
        eng = _run(cfg, met, **kw)
        T = eng.state_numpy()
        for k in STATE:
            assert np.array_equal(S[k], T[k]), (kw, k)
        assert np.array_equal(ref.escaped.cpu().numpy(), eng.escaped.cpu().numpy()), kw
        assert np.array_equal(ref.active_steps.cpu().numpy(), eng.active_steps.cpu().numpy()), kw
        a, b = ref.footprint.cpu().numpy(), eng.footprint.cpu().numpy()
        if "reorder_every" in kw:
            np.testing.assert_allclose(b, a, rtol=2e-5, atol=0)
        else:
            assert np.array_equal(a, b), kw


def test_tuner_picks_a_candidate_and_caches_it(setup, tmp_path, monkeypatch):
    cfg, met = setup
    monkeypatch.setenv("GLIDE_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    cands = [dict(step_block=128, reorder_every=0), dict(step_block=256, reorder_every=0)]
    first = tune(cfg, met, candidates=cands, steps=8, compact=True)
    assert not first["cached"] and first["best"] in cands and len(first["trials"]) == 2
    assert all(t["particle_steps_per_s"] > 0 for t in first["trials"])
    again = tune(cfg, met, candidates=cands, steps=8, compact=True)
    assert again["cached"] and again["best"] == first["best"]
    assert (tmp_path / "tune.json").exists()
