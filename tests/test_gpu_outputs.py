"""GPU: trajectory diagnostics and batch outputs through the C ABI.

* glide_particle_means / glide_field_means: fixed-order float64 reductions,
  against numpy's float64 mean of the same float32 values (relative 1e-12: the
  summation orders differ, each in float64), bit-identical between two runs and
  between eager steps and graph replay (the captured body reads its output row
  from the device cursor).
* trajectory rows: active and escaped per step against the float64 oracle's
  (active exactly: the window masks are schedule arithmetic; escape totals
  exactly, per step within one), alive = N - cumulative escapes.
* write_outputs: the footprint store, endpoints and trajectory tables hold
  exactly the engine's arrays."""

from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd.output import ZarrFootprintStore  # noqa: E402
from glide_amd.runner import BatchEngine  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402
from oracle import run as ORUN  # noqa: E402


@pytest.fixture(scope="module")
def setup():
    cfg = workload_config("batched-cpu-6r")
    met = SyntheticMet(cfg["spec"])
    return cfg, met


def _host_mean(eng) -> np.ndarray:
    S = eng.endpoints()
    return np.array([S[k].astype(np.float64).mean() for k in ("lon", "lat", "alt")])


def test_position_means_eager_and_graph(setup):
    cfg, met = setup
    eng = BatchEngine(cfg, device="cuda", met=met, diagnostics=True, compact=True)
    checkpoints = [1, 7, len(eng.sched) // 2, len(eng.sched)]
    done = 0
    for k in checkpoints:
        eng.run(k - done)
        done = k
        got = eng.pos_means[k - 1].cpu().numpy()
        np.testing.assert_allclose(got, _host_mean(eng), rtol=1e-12, atol=0)
    eager = eng.pos_means.cpu().numpy().copy()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = BatchEngine(cfg, device="cuda", met=met, diagnostics=True, compact=True, use_graph=True)
        g.run()
    torch.cuda.synchronize()
    assert g.graph_launches > 0
    assert np.array_equal(g.pos_means.cpu().numpy(), eager)
    again = BatchEngine(cfg, device="cuda", met=met, diagnostics=True, compact=True)
    again.run()
    assert np.array_equal(again.pos_means.cpu().numpy(), eager)
    g.close()


def test_field_means_match_numpy(setup):
    cfg, met = setup
    eng = BatchEngine(cfg, device="cuda", met=met, diagnostics=True)
    eng.run(3)
    torch.cuda.synchronize()
    assert eng.hour_wind
    for h, m in eng.hour_wind.items():
        out3 = eng.met.resampled(h, 0)[0].cpu().numpy().astype(np.float64)
        ref = out3[:3].reshape(3, -1).mean(axis=1)
        np.testing.assert_allclose(m.cpu().numpy(), ref, rtol=1e-12, atol=1e-15)


def test_trajectory_rows_and_outputs(setup, tmp_path):
    cfg, met = setup
    eng = BatchEngine(cfg, device="cuda", met=met, diagnostics=True, compact=True)
    eng.run()
    rows = eng.trajectory_rows(batch_idx=0)
    ref = ORUN.run(cfg, dty=np.float64, met=met)
    assert len(rows) == len(eng.sched) == len(ref.active_per_step)
    act = [r["active_particles"] for r in rows]
    esc = [r["escaped_this_step"] for r in rows]
    assert act == ref.active_per_step
    # float32 and float64 trajectories may cross the domain edge a step apart
    assert sum(esc) == sum(ref.escaped_per_step)
    assert max(abs(a - b) for a, b in zip(esc, ref.escaped_per_step)) <= 1
    assert [r["alive_particles"] for r in rows] == list(eng.n - np.cumsum(esc))
    assert [r["step"] for r in rows] == list(range(1, len(rows) + 1))
    st0 = eng.sched[0]
    w0 = eng.hour_wind[st0.bracket_hour].cpu().numpy() * (1 - st0.alpha) + \
        eng.hour_wind[st0.bracket_hour + 1].cpu().numpy() * st0.alpha
    assert np.allclose([rows[0]["u_mean_m_s"], rows[0]["v_mean_m_s"], rows[0]["w_mean_m_s"]], w0, rtol=1e-15)
    paths = eng.write_outputs(tmp_path)
    fp = eng.footprint_numpy()
    st = ZarrFootprintStore(paths["footprint"])
    assert st.shape() == fp.shape
    for r in range(fp.shape[0]):
        assert np.array_equal(st.read(r), fp[r])
    assert np.array_equal(st.read_coord("release_lon"), [rel.lon for rel in cfg["releases"]])
    ep = pd.read_parquet(paths["endpoints"])
    S = eng.endpoints()
    assert list(ep.columns) == ["lon", "lat", "alt", "weight", "release_idx"]
    for k in ("lon", "lat", "alt", "weight", "release_idx"):
        assert np.array_equal(ep[k].to_numpy(), S[k]), k
    tr = pd.read_parquet(paths["trajectory"])
    assert len(tr) == len(rows) and list(tr["elapsed_seconds"]) == [r["step"] * cfg["dt"] for r in rows]
    np.testing.assert_array_equal(tr["mean_lat"].to_numpy(), [r["mean_lat"] for r in rows])
