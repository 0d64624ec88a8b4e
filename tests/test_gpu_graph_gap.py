"""GPU: the captured step body across met hours with no live particle.

Two releases a few hours apart with one-hour backward windows, convection
off, hourly compaction and graph replay: between the later release's terminus
and the earlier one's window end no particle can be active, the compaction
leaves an empty live set, and the captured glide_step_table runs over no
particles. Every such call must still advance the device cursor once, so the
replays that follow read their own table rows. The graph run must equal the
eager run bit for bit (particles, footprint, totals, per-step counts)."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd.runner import BatchEngine  # noqa: E402
from glide_amd.synthetic import MetSpec, Release, SyntheticMet, _domain_and_grid  # noqa: E402

STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "alive")


def _gap_config():
    spec = MetSpec(nx=64, ny=48, za=40, n_hours=8, seed=20240111)
    lon = float(spec.lon[30])
    lat = float(spec.lat[20])
    rels = [
        Release(lon=lon, lat=lat, alt_agl_m=50.0, n_particles=3000, release_time_s=2 * 3600 - 120,
                length_s=3600, seed=31, release_idx=0),
        Release(lon=lon + 0.5, lat=lat - 0.5, alt_agl_m=120.0, n_particles=5000, release_time_s=6 * 3600 - 120,
                length_s=3600, seed=32, release_idx=1),
    ]
    grid, domain = _domain_and_grid(spec, len(rels), 2)
    return dict(spec=spec, releases=rels, grid=grid, domain=domain, convection=False, dt=60, seed=17)


@pytest.fixture(scope="module")
def cfg_met():
    cfg = _gap_config()
    return cfg, SyntheticMet(cfg["spec"])


@pytest.mark.parametrize("accumulate", ["fixed", "sort"])
def test_graph_replay_across_empty_hours_equals_eager(cfg_met, accumulate):
    cfg, met = cfg_met
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eager = BatchEngine(cfg, device="cuda", met=met, compact=True, use_graph=False, accumulate=accumulate)
        eager.run()
        graph = BatchEngine(cfg, device="cuda", met=met, compact=True, use_graph=True, accumulate=accumulate)
        live = []
        while graph.step_i < len(graph.sched):
            graph.run_step()
            live.append(graph.n_live)
        graph.join()
    torch.cuda.synchronize()
    assert min(live) == 0, "the schedule must hold hours with no live particle"
    assert graph.graph_launches > 0
    assert int(graph.row[0].item()) == len(graph.sched), "every step advanced the device cursor once"
    a, b = eager.state_numpy(), graph.state_numpy()
    for name in STATE:
        assert np.array_equal(a[name], b[name]), name
    assert np.array_equal(eager.footprint.cpu().numpy(), graph.footprint.cpu().numpy())
    assert np.array_equal(eager.release_total.cpu().numpy(), graph.release_total.cpu().numpy())
    assert np.array_equal(eager.escaped.cpu().numpy(), graph.escaped.cpu().numpy())
    assert np.array_equal(eager.active_steps.cpu().numpy(), graph.active_steps.cpu().numpy())
    # both releases were stepped for their whole windows
    act = eager.active_steps.cpu().numpy()
    assert act.sum() > 0 and (act == 0).sum() > 0
    graph.close()


@pytest.mark.parametrize("accumulate", ["fixed", "sort"])
def test_multi_step_graphs_equal_eager(cfg_met, accumulate):
    """run() with use_graph launches runs of replayable steps as graphs of
    16, 8, 4, 2 and 1 captured steps (re-binning every 5 steps and the hourly
    compaction break the runs): the same particles, footprint, totals and
    per-step counts as the eager run, bit for bit, in far fewer launches."""
    cfg, met = cfg_met
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eager = BatchEngine(cfg, device="cuda", met=met, compact=True, reorder_every=5, accumulate=accumulate)
        eager.run()
        graph = BatchEngine(cfg, device="cuda", met=met, compact=True, reorder_every=5, use_graph=True,
                            accumulate=accumulate)
        graph.run(37)  # a run that ends inside a replayable stretch
        graph.run()
    torch.cuda.synchronize()
    n = len(graph.sched)
    sizes = sorted({k[2] for k in graph._graphs})
    print(f"{n} steps in {graph.graph_launches} graph launches, graph sizes {sizes}", flush=True)
    assert graph.step_i == n and int(graph.row[0].item()) == n
    assert max(sizes) > 1 and graph.graph_launches < n // 2
    a, b = eager.state_numpy(), graph.state_numpy()
    for name in STATE:
        assert np.array_equal(a[name], b[name]), name
    assert np.array_equal(eager.footprint.cpu().numpy(), graph.footprint.cpu().numpy())
    assert np.array_equal(eager.release_total.cpu().numpy(), graph.release_total.cpu().numpy())
    assert np.array_equal(eager.escaped.cpu().numpy(), graph.escaped.cpu().numpy())
    assert np.array_equal(eager.active_steps.cpu().numpy(), graph.active_steps.cpu().numpy())
    graph.close()
