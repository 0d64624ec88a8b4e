"""GPU: continuous-batching serving. Requests admitted at different times and
stepped together, each on its own cursor, give the same result as when each
runs alone, and the same as the single-release engine (pinned to the oracle
by tests/test_gpu_parity.py) on the same release: every particle
bit-identical; with the default fixed-point footprint (exact integer sums)
the footprints and totals of a request alone and in a mix are bit-identical
too, and match the engine's within float32 rounding (2e-5); the sort
accumulation's float32 sums match within summation-order rounding (2e-5),
totals within 1e-9."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd import _lib as L  # noqa: E402
from glide_amd.runner import BatchEngine  # noqa: E402
from glide_amd.serving import MetRing, Request, Server  # noqa: E402
from glide_amd.synthetic import Release, SyntheticMet, workload_config  # noqa: E402

STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "alive")
SEED = 13


@pytest.fixture(scope="module")
def setup():
    cfg = workload_config("batched-cpu-6r")  # 64x48x40 grid, met hours 0..6
    met = SyntheticMet(cfg["spec"])
    ring = MetRing(met, first_hour=0, n_hours=6, met_dtype=L.GLIDE_MET_BF16)
    lon, lat = met.lon, met.lat
    reqs = [
        Request(req_id=101, lon=float(lon[20]) + 0.07, lat=float(lat[20]) - 0.03, alt_agl_m=40.0, n_particles=1500,
                release_time_s=3 * 3600 + 900, length_s=2 * 3600, seed=5),
        Request(req_id=102, lon=float(lon[40]) - 0.11, lat=float(lat[30]) + 0.02, alt_agl_m=150.0, n_particles=700,
                release_time_s=4 * 3600 + 300, length_s=3600, seed=6),
        Request(req_id=103, lon=float(lon[33]), lat=float(lat[12]), alt_agl_m=10.0, n_particles=2500,
                release_time_s=5 * 3600 - 120, length_s=2 * 3600 + 1800, seed=7),
        Request(req_id=104, lon=float(lon[10]) + 0.2, lat=float(lat[40]), alt_agl_m=80.0, n_particles=1000,
                release_time_s=2 * 3600 + 2400, length_s=3600 + 600, seed=8),
    ]
    return cfg, met, ring, reqs


def _server(cfg, ring, slots=8):
    return Server(ring, cfg["grid"], cfg["domain"], max_slots=slots, pool_particles=8192, n_t_max=3, seed=SEED,
                  keep_state=True)


def _by_uid(state):
    o = np.argsort(state["uid"].view(np.uint32), kind="stable")
    return {k: v[o] for k, v in state.items()}


@pytest.fixture(scope="module")
def mixed(setup):
    cfg, met, ring, reqs = setup
    srv = _server(cfg, ring)
    assert srv.admit(reqs[0]) and srv.admit(reqs[1])
    for _ in range(13):
        srv.step()
    assert srv.admit(reqs[2])
    for _ in range(29):
        srv.step()
    assert srv.admit(reqs[3])
    out = {r.req.req_id: r for r in srv.drain()}
    assert sorted(out) == [r.req_id for r in reqs]
    return out


def test_request_alone_equals_request_in_mix(setup, mixed):
    cfg, met, ring, reqs = setup
    for req in reqs:
        srv = _server(cfg, ring)
        assert srv.admit(req)
        (alone,) = srv.drain()
        mix = mixed[req.req_id]
        a, b = _by_uid(alone.state), _by_uid(mix.state)
        for k in STATE:
            assert np.array_equal(a[k], b[k]), (req.req_id, k)
        assert alone.steps == mix.steps == int(np.ceil(req.length_s / 60))
        assert alone.escaped == mix.escaped and alone.active_steps == mix.active_steps
        assert np.array_equal(alone.footprint != 0, mix.footprint != 0)
        np.testing.assert_allclose(mix.footprint, alone.footprint, rtol=2e-5, atol=0)
        assert abs(mix.total - alone.total) <= 1e-9 * alone.total
        assert alone.total > 0
        if srv.fixed_mode:  # exact sums: the order of the atomics cannot show
            assert np.array_equal(mix.footprint, alone.footprint) and mix.total == alone.total


def test_served_request_equals_single_release_engine(setup, mixed):
    """The same release through BatchEngine (a batch of one, default mask, the
    engine the oracle pins): identical particles, matching footprint."""
    cfg, met, ring, reqs = setup
    req = reqs[0]
    rel = Release(lon=req.lon, lat=req.lat, alt_agl_m=req.alt_agl_m, n_particles=req.n_particles,
                  release_time_s=req.release_time_s, duration_s=req.duration_s, length_s=req.length_s,
                  seed=req.seed, release_idx=0, noise_id=req.req_id)
    n_t = int(np.ceil(req.length_s / 3600))
    c1 = dict(cfg, releases=[rel], grid=dict(cfg["grid"], n_r=1, n_t=n_t), seed=SEED, convection=False)
    eng = BatchEngine(c1, device="cuda", met=met, met_dtype=L.GLIDE_MET_BF16)
    eng.run()
    S = eng.state_numpy()
    b = _by_uid(mixed[req.req_id].state)
    for k in STATE:
        assert np.array_equal(S[k], b[k]), k
    fe = eng.footprint.cpu().numpy().reshape(mixed[req.req_id].footprint.shape)
    np.testing.assert_allclose(mixed[req.req_id].footprint, fe, rtol=2e-5, atol=0)


def test_admission_waits_for_free_slots(setup):
    cfg, met, ring, reqs = setup
    srv = _server(cfg, ring, slots=1)
    assert srv.admit(reqs[1])
    assert not srv.admit(reqs[0])  # the only slot is busy
    out = srv.drain()
    assert [r.req.req_id for r in out] == [reqs[1].req_id]
    assert srv.admit(reqs[0])  # free again
    assert [r.req.req_id for r in srv.drain()] == [reqs[0].req_id]


def test_finished_footprints_stream_to_zarr(setup, mixed, tmp_path):
    """With a ResultWriter, every finished request lands in its own Zarr store
    (zlib chunks) holding exactly the footprint the server produced, with the
    request's coordinates; results count as done only once written."""
    from glide_amd.output import ZarrFootprintStore
    from glide_amd.serving import ResultWriter

    cfg, met, ring, reqs = setup
    writer = ResultWriter(tmp_path, workers=2)
    srv = Server(ring, cfg["grid"], cfg["domain"], max_slots=8, pool_particles=8192, n_t_max=3, seed=SEED,
                 writer=writer)
    for r in reqs:
        assert srv.admit(r)
    out = {r.req.req_id: r for r in srv.drain()}
    writer.close()
    assert sorted(out) == [r.req_id for r in reqs] and not writer.busy
    assert len(writer.paths) == len(reqs)
    for rid, res in out.items():
        st = ZarrFootprintStore(tmp_path / f"request_{rid:06d}" / "footprint.zarr")
        assert st.shape() == (1,) + res.footprint.shape
        assert np.array_equal(st.read(0), res.footprint)
        np.testing.assert_allclose(res.footprint, mixed[rid].footprint, rtol=2e-5, atol=0)
        assert st.read_coord("release_lon")[0] == pytest.approx(res.req.lon)
        assert np.isfinite(res.latency_s)


def test_sort_accumulation_server_matches_fixed(setup, mixed):
    """A server with the sort accumulation (keys and values, radix sort,
    fixed-order float32 sums) gives the same particles and the fixed-point
    server's footprint within float32 summation rounding."""
    cfg, met, ring, reqs = setup
    srv = Server(ring, cfg["grid"], cfg["domain"], max_slots=8, pool_particles=8192, n_t_max=3, seed=SEED,
                 keep_state=True, accumulate="sort")
    assert not srv.fixed_mode
    req = reqs[0]
    assert srv.admit(req)
    (res,) = srv.drain()
    ref = mixed[req.req_id]
    a, b = _by_uid(res.state), _by_uid(ref.state)
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(res.footprint != 0, ref.footprint != 0)
    np.testing.assert_allclose(res.footprint, ref.footprint, rtol=2e-5, atol=0)
    assert abs(res.total - ref.total) <= 1e-9 * ref.total


def test_out_of_range_wave_slots_mark_unused_blocks(setup):
    """wave_slot entries outside [0, max_slots) (the footprint's n_r, the
    size of the slots array) mark unused blocks like -1: a server whose free
    blocks hold such values on the device serves a request bit for bit like
    an untouched server."""
    cfg, met, ring, reqs = setup
    out = []
    for poke in (False, True):
        srv = _server(cfg, ring)
        assert srv.admit(reqs[0])
        if poke:
            free = np.flatnonzero(srv.wave_slot_h < 0)
            assert free.size > 3
            bad = np.resize(np.array([srv.max_slots, srv.max_slots + 1000, 2**31 - 1], np.int32), free.size)
            dev = srv.wave_slot.device
            srv.wave_slot[torch.from_numpy(free).to(dev)] = torch.from_numpy(bad).to(dev)
            torch.cuda.synchronize()
            print(f"wave_slot: {free.size} free blocks set to {bad[:3].tolist()}", flush=True)
        (res,) = srv.drain()
        out.append(res)
    a, b = _by_uid(out[0].state), _by_uid(out[1].state)
    for k in STATE:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(out[0].footprint, out[1].footprint) and out[0].total == out[1].total
    assert out[0].escaped == out[1].escaped and out[0].active_steps == out[1].active_steps
    assert out[0].total > 0
