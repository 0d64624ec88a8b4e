"""GPU: the benchmark workloads at their full sizes, through properties that do
not need the oracle (which cannot run 2^20 particles in test time):

* plain FP32 workload (2^20 particles, 240x160x60 met, per-column convection):
  two runs of the first two met hours are bit-identical (determinism of the
  sort-based accumulation at full size); eager steps and graph replay are
  bit-identical; every step's active count equals the schedule's window count;
  the float32 footprint sums (in float64) to the float64 release total within
  float32 rounding (mass conservation of the accumulation);
* batched bfloat16 workload (336 releases, 2.4 M particles): per-release slab
  sums equal the float64 release totals, and compaction on/off gives the same
  particles.

The engines use the default fixed-point accumulation; test_gpu_fixed.py pins it
bit for bit against exact integer sums at small size."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd import _lib as L  # noqa: E402
from glide_amd.runner import BatchEngine, DeviceMet  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402

STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "alive")
PLAIN_STEPS = 120  # two met hours, one convection event


@pytest.fixture(scope="module")
def plain():
    cfg = workload_config("plain-gpu-1M-24h")
    met = SyntheticMet(cfg["spec"])
    print("building plain-gpu-1M-24h engine", flush=True)
    a = BatchEngine(cfg, device="cuda", met=met)
    a.run(PLAIN_STEPS)
    torch.cuda.synchronize()
    return cfg, met, a


def test_plain_full_size_deterministic_and_conserving(plain):
    cfg, met, a = plain
    b = BatchEngine(cfg, device="cuda", met=met, device_met=a.met)
    b.run(PLAIN_STEPS)
    torch.cuda.synchronize()
    sa, sb = a.state_numpy(), b.state_numpy()
    for k in STATE:
        assert np.array_equal(sa[k], sb[k]), k
    fa = a.footprint.cpu().numpy()
    assert np.array_equal(fa, b.footprint.cpu().numpy())
    tot = float(a.release_total.cpu().numpy()[0])
    assert tot > 0
    assert abs(fa.astype(np.float64).sum() - tot) <= 2e-6 * tot
    act = a.active_steps.cpu().numpy()[:PLAIN_STEPS]
    want = [a.active_count(st) for st in a.sched[:PLAIN_STEPS]]
    esc = np.cumsum(a.escaped.cpu().numpy()[:PLAIN_STEPS])
    # active = in the release window and not escaped before this step
    assert np.array_equal(act, np.array(want) - np.concatenate([[0], esc[:-1]]))
    assert act.max() == a.n


def test_plain_full_size_graph_replay_matches_eager(plain):
    cfg, met, a = plain
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = BatchEngine(cfg, device="cuda", met=met, device_met=a.met, use_graph=True, overlap=True)
        g.run(PLAIN_STEPS)
    torch.cuda.synchronize()
    assert g.graph_launches > 0
    sa, sg = a.state_numpy(), g.state_numpy()
    for k in STATE:
        assert np.array_equal(sa[k], sg[k]), k
    assert np.array_equal(a.footprint.cpu().numpy(), g.footprint.cpu().numpy())
    assert np.array_equal(a.escaped.cpu().numpy(), g.escaped.cpu().numpy())
    g.close()


def test_batched_full_size_totals_and_compaction():
    cfg = workload_config("batched-gpu-336r")
    met = SyntheticMet(cfg["spec"])
    hours = list(range(cfg["spec"].n_hours))
    print("building batched-gpu-336r met", flush=True)
    dm = DeviceMet(met, hours, met_dtype=L.GLIDE_MET_BF16)
    steps = 1600  # 27 met hours: past the terminus of the first 24 h releases
    kw = dict(met_dtype=L.GLIDE_MET_BF16, compensated=True, device_met=dm)
    a = BatchEngine(cfg, device="cuda", met=met, compact=True, overlap=True, **kw)
    a.run(steps)
    b = BatchEngine(cfg, device="cuda", met=met, compact=False, **kw)
    b.run(steps)
    torch.cuda.synchronize()
    assert a.n_live < a.n  # compaction dropped particles that can never be active again
    sa, sb = a.state_numpy(), b.state_numpy()
    for k in STATE:
        assert np.array_equal(sa[k], sb[k]), k
    R = cfg["grid"]["n_r"]
    slabs = a.footprint.cpu().numpy().astype(np.float64)
    if a.comp is not None:  # Neumaier terms of the sort accumulation
        slabs = slabs + a.comp.cpu().numpy().astype(np.float64)
    slabs = slabs.reshape(R, -1)
    tot = a.release_total.cpu().numpy()
    live = tot > 0
    assert live.sum() > 10
    rel = np.abs(slabs.sum(1)[live] - tot[live]) / tot[live]
    print("releases with residence", int(live.sum()), "max relative slab-vs-total", rel.max())
    assert rel.max() <= 1e-6
    assert np.all(slabs[~live] == 0)
