"""GPU: the precision layout of positions (GlideParticles.cell): each
particle's horizontal position as its nearest met grid point (int32) plus
float32 offsets in degrees, against the float64 oracle and against the
absolute float32 layout.

* Per step over the first met hour (10k particles, 3 h workload): each engine
  step from the float64 state the layout holds, against one float64 oracle
  step: at least 99.9 % of particles within 1e-5 relative (the bar of the
  absolute layout), and the mean relative lon/lat error below the absolute
  layout's (whose float32 coordinates round at ~3e-8 of their magnitude).
* Full run: footprint within 1 % relative L2 of the float64 oracle and
  escapes equal.
* The layout keeps every offset within 0.6 of a cell and every cell on the
  grid; re-binning and hourly compaction leave a batch's results bit for bit
  unchanged, and the captured step replays bit for bit like eager steps."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd.runner import BatchEngine  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402
from oracle import lpdm as O  # noqa: E402
from oracle import resample as ORS  # noqa: E402
from oracle import run as ORUN  # noqa: E402


@pytest.fixture(scope="module")
def cfg0():
    cfg = workload_config("plain-cpu-10k-3h")
    return cfg, SyntheticMet(cfg["spec"])


def _per_step_errors(cfg, met, cell, n_steps=61):
    eng = BatchEngine(cfg, device="cuda", met=met, cell_positions=cell)
    ra = eng.ra
    hours = {}

    def hour(h):
        if h not in hours:
            hours[h] = ORS.resample_source_hour(met.hour(h), met, met.spec.agl_levels)
        return hours[h]

    windows = {}
    S = eng.state_numpy()
    p = O.HannaParams()
    errs, fracs = [], []
    for st in eng.sched[:n_steps]:
        bh = st.bracket_hour
        if bh not in windows:
            windows[bh] = ORUN.build_window(hour(bh), hour(bh + 1), met.spec.agl_levels, met.lon, met.lat,
                                            np.float64)
        active = (ra.r_off >= np.float32(st.cursor_off)) & (ra.lo_off <= np.float32(st.cursor_off))
        pre = {k: np.asarray(S[k], np.float64) for k in ("lon", "lat", "z", "up", "vp", "wp", "um", "vm")}
        pre["uid"] = S["uid"]
        eng.run_step()
        S = eng.state_numpy()
        ref = O.step(pre, active, windows[bh], st.alpha, float(st.delta_s), p, seed=eng.seed,
                     step_index=st.noise_step)
        if not active.any():
            continue
        e_lon = np.abs(S["lon"][active] - ref["lon"][active]) / np.abs(ref["lon"][active])
        e_lat = np.abs(S["lat"][active] - ref["lat"][active]) / np.abs(ref["lat"][active])
        zs = np.maximum(np.maximum(np.abs(ref["z"][active]), np.abs(pre["z"][active])), 10.0)
        e_z = np.abs(S["z"][active] - ref["z"][active]) / zs
        errs.append((e_lon.mean(), e_lat.mean(), e_z.mean()))
        fracs.append(float((np.maximum(np.maximum(e_lon, e_lat), e_z) <= 1e-5).mean()))
    return eng, np.array(errs).mean(axis=0), min(fracs)


def test_cell_layout_per_step_errors_against_float64(cfg0):
    cfg, met = cfg0
    eng, e_cell, f_cell = _per_step_errors(cfg, met, True)
    _, e_abs, f_abs = _per_step_errors(cfg, met, False)
    print(f"mean relative error lon/lat/z: cell {e_cell}, absolute {e_abs}; min fraction within 1e-5: "
          f"cell {f_cell:.5f}, absolute {f_abs:.5f}")
    assert f_cell >= 0.999 and f_abs >= 0.999
    assert e_cell[0] < e_abs[0] and e_cell[1] < e_abs[1]
    # offsets within 0.6 of a cell, cells on the grid
    b = eng.bufs[eng.cur]
    g = eng.met.grid
    cell = b["cell"].cpu().numpy()
    assert (cell & 0xFFFF).max() < g.nx and (cell >> 16).max() < g.ny and cell.min() >= 0
    dlon, dlat = (g.lon1 - g.lon0) / (g.nx - 1), abs(g.lat1 - g.lat0) / (g.ny - 1)
    alive = b["alive"].cpu().numpy() != 0
    assert np.abs(b["lon"].cpu().numpy()[alive]).max() <= 0.6 * dlon
    assert np.abs(b["lat"].cpu().numpy()[alive]).max() <= 0.6 * dlat


def test_cell_layout_full_run_footprint(cfg0):
    cfg, met = cfg0
    eng = BatchEngine(cfg, device="cuda", met=met, cell_positions=True)
    eng.run()
    torch.cuda.synchronize()
    fp = eng.footprint.cpu().numpy().astype(np.float64)
    ref = ORUN.run(cfg, dty=np.float64, met=met)
    l2 = np.linalg.norm(fp - ref.footprint) / np.linalg.norm(ref.footprint)
    tot = abs(float(eng.release_total.cpu().numpy().sum()) - ref.footprint.sum()) / ref.footprint.sum()
    print(f"cell layout full run: L2 {l2:.2e}, total {tot:.2e}")
    assert l2 <= 0.01 and tot <= 1e-3
    assert int(eng.escaped.cpu().numpy().sum()) == int(sum(ref.escaped_per_step))


def test_cell_layout_order_free_and_graph_replay():
    cfg = workload_config("batched-cpu-6r")
    met = SyntheticMet(cfg["spec"])
    runs = {}
    stream = torch.cuda.Stream()  # graph capture needs a non-default stream
    with torch.cuda.stream(stream):
        for key, kw in (("plain", {}), ("binned", dict(compact=True, reorder_every=8)),
                        ("graph", dict(compact=True, use_graph=True))):
            eng = BatchEngine(cfg, device="cuda", met=met, cell_positions=True, **kw)
            eng.run()
            torch.cuda.synchronize()
            runs[key] = (eng.state_numpy(), eng.footprint.cpu().numpy(), eng.escaped.cpu().numpy())
            if key == "graph":
                assert eng.graph_launches > 0
    a = runs["plain"]
    for key in ("binned", "graph"):
        b = runs[key]
        for k in ("lon", "lat", "z", "up", "vp", "wp", "alive"):
            assert np.array_equal(a[0][k], b[0][k]), (key, k)
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), key
