"""GPU: continuous-batching serving. Requests admitted at different times and
stepped together, each on its own cursor, give the same result as when each
runs alone, and the same as the single-release engine (pinned to the oracle
by tests/test_gpu_parity.py) on the same release: every particle
bit-identical, and with the fixed-point footprint (exact integer sums, the
same scale) footprints and totals bit-identical too; with per-slot convection
as well. Against the float64 oracle directly, served requests on the
bfloat16 ring are within the batched method's bfloat16 bars (footprint
relative L2 <= 1 %, residence total <= 1e-2). Admission follows free HBM: the
pool and the footprint arena grow when a request does not fit (results
unchanged), and a request that cannot fit within the server's memory limit
waits."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd import _lib as L  # noqa: E402
from glide_amd.runner import BatchEngine, ConvectionOptions  # noqa: E402
from glide_amd.serving import BLOCK, MetRing, Request, Server  # noqa: E402
from glide_amd.synthetic import Release, SyntheticMet, workload_config  # noqa: E402

STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "alive")
SEED = 13


@pytest.fixture(scope="module")
def setup():
    cfg = workload_config("batched-cpu-6r")  # 64x48x40 grid, met hours 0..6
    met = SyntheticMet(cfg["spec"])
    ring = MetRing(met, first_hour=0, n_hours=6, met_dtype=L.GLIDE_MET_BF16, convection=True)
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


def _server(cfg, ring, slots=8, conv=False, **kw):
    return Server(ring, cfg["grid"], cfg["domain"], max_slots=slots, pool_particles=kw.pop("pool", 8192),
                  n_t_max=3, seed=SEED, keep_state=True, convection=ConvectionOptions() if conv else None, **kw)


def _by_uid(state):
    o = np.argsort(state["uid"].view(np.uint32), kind="stable")
    return {k: v[o] for k, v in state.items()}


def _mix(cfg, ring, reqs, conv):
    srv = _server(cfg, ring, conv=conv)
    assert srv.admit(reqs[0]) and srv.admit(reqs[1])
    for _ in range(13):
        srv.step()
    assert srv.admit(reqs[2])
    for _ in range(29):
        srv.step()
    assert srv.admit(reqs[3])
    out = {r.req.req_id: r for r in srv.drain()}
    assert sorted(out) == [r.req_id for r in reqs]
    if conv:
        assert srv.convect_launches > 0
    return out


@pytest.fixture(scope="module")
def mixed(setup):
    cfg, met, ring, reqs = setup
    return _mix(cfg, ring, reqs, conv=False)


@pytest.fixture(scope="module")
def mixed_conv(setup):
    cfg, met, ring, reqs = setup
    return _mix(cfg, ring, reqs, conv=True)


def _same(a_res, b_res, tag):
    a, b = _by_uid(a_res.state), _by_uid(b_res.state)
    for k in STATE:
        assert np.array_equal(a[k], b[k]), (tag, k)
    assert a_res.escaped == b_res.escaped and a_res.active_steps == b_res.active_steps, tag
    assert np.array_equal(a_res.footprint, b_res.footprint) and a_res.total == b_res.total, tag


@pytest.mark.parametrize("conv", [False, True])
def test_request_alone_equals_request_in_mix(setup, mixed, mixed_conv, conv):
    cfg, met, ring, reqs = setup
    mix_all = mixed_conv if conv else mixed
    for req in reqs:
        srv = _server(cfg, ring, conv=conv)
        assert srv.admit(req)
        (alone,) = srv.drain()
        mix = mix_all[req.req_id]
        _same(alone, mix, req.req_id)
        assert alone.steps == mix.steps == int(np.ceil(req.length_s / 60))
        assert alone.total > 0


def _alone_cfg(cfg, req, conv):
    """The batch holding the request's release alone (its particle ids, the server's seed)."""
    rel = Release(lon=req.lon, lat=req.lat, alt_agl_m=req.alt_agl_m, n_particles=req.n_particles,
                  release_time_s=req.release_time_s, duration_s=req.duration_s, length_s=req.length_s,
                  seed=req.seed, release_idx=0, noise_id=req.req_id)
    n_t = int(np.ceil(req.length_s / 3600))
    return dict(cfg, releases=[rel], grid=dict(cfg["grid"], n_r=1, n_t=n_t), seed=SEED, convection=conv)


@pytest.mark.parametrize("conv", [False, True])
def test_served_request_equals_single_release_engine(setup, mixed, mixed_conv, conv):
    """The same release through BatchEngine (a batch of one, the engine the
    oracle pins), convection on or off: identical particles, and with the same
    fixed-point scale an identical footprint and total."""
    cfg, met, ring, reqs = setup
    for req in reqs[:2]:
        eng = BatchEngine(_alone_cfg(cfg, req, conv), device="cuda", met=met, met_dtype=L.GLIDE_MET_BF16)
        eng.run()
        S = eng.state_numpy()
        res = (mixed_conv if conv else mixed)[req.req_id]
        b = _by_uid(res.state)
        for k in STATE:
            assert np.array_equal(S[k], b[k]), (req.req_id, k)
        fe = eng.footprint.cpu().numpy().reshape(res.footprint.shape)
        assert np.array_equal(res.footprint, fe), req.req_id
        assert res.total == float(eng.release_total.cpu().numpy()[0])
        assert res.escaped == int(eng.escaped.cpu().numpy().sum())


def test_convection_moves_served_particles(setup, mixed, mixed_conv):
    """Convection is on in the convective server: a request's particles end
    elsewhere than without it (a request released under firing columns)."""
    cfg, met, ring, reqs = setup
    moved = [not np.array_equal(_by_uid(mixed[r.req_id].state)["z"], _by_uid(mixed_conv[r.req_id].state)["z"])
             for r in reqs]
    assert any(moved), moved


def test_served_requests_match_the_float64_oracle(setup, mixed_conv):
    """Served requests with convection on the bfloat16 ring against the
    float64 oracle run of the release alone (its own terminus, the same
    Philox draws, per-column convection): within the batched method's
    bfloat16 bars."""
    from oracle import run as ORUN

    cfg, met, ring, reqs = setup
    for req in reqs:
        c1 = _alone_cfg(cfg, req, True)
        hours = {}
        ref = ORUN.run(c1, np.float64, met=met, hours_cache=hours, release_terminus=True,
                       convection=ORUN.column_convection(met, hours, SEED))
        got = mixed_conv[req.req_id]
        rf = ref.footprint.reshape(got.footprint.shape)
        l2 = np.linalg.norm(got.footprint - rf) / np.linalg.norm(rf)
        tot = abs(got.total - rf.sum()) / rf.sum()
        print(f"request {req.req_id}: footprint L2 {l2:.2e}, total {tot:.2e}")
        assert l2 <= 0.01 and tot <= 1e-2, (req.req_id, l2, tot)
        # bfloat16 winds move a few particles across the domain edge at another step
        assert abs(got.escaped - int(sum(ref.escaped_per_step))) <= max(2, req.n_particles // 100)


def test_admission_waits_for_free_slots(setup):
    cfg, met, ring, reqs = setup
    srv = _server(cfg, ring, slots=1)
    assert srv.admit(reqs[1])
    assert not srv.admit(reqs[0])  # the only slot is busy
    out = srv.drain()
    assert [r.req.req_id for r in out] == [reqs[1].req_id]
    assert srv.admit(reqs[0])  # free again
    assert [r.req.req_id for r in srv.drain()] == [reqs[0].req_id]


def test_admission_grows_pool_and_arena_within_free_hbm(setup, mixed):
    """A server whose pool and arena start smaller than one request grows
    them on admission (the requests' results bit-identical to the roomy
    server's); one whose memory limit leaves no room makes a request wait until
    a running one finishes and frees its blocks and slab."""
    cfg, met, ring, reqs = setup
    srv = _server(cfg, ring, pool=64, arena_cells=1)
    b0, c0 = srv.n_blocks, srv.arena.size
    for r in reqs:
        assert srv.admit(r)
    assert srv.n_blocks > b0 and srv.arena.size > c0
    out = {r.req.req_id: r for r in srv.drain()}
    for r in reqs:
        _same(out[r.req_id], mixed[r.req_id], r.req_id)
    # limited: room for the first request's particles and slab, not for a second beside it
    small = _server(cfg, ring, pool=64, arena_cells=1)
    lim = small.device_bytes()
    nb = (reqs[0].n_particles + BLOCK - 1) // BLOCK
    need = nb * BLOCK * small.bytes_per_particle() + 2 * small.cpb * 8
    tight = _server(cfg, ring, pool=64, arena_cells=1, hbm_limit_bytes=int(lim + need * 1.2))
    assert tight.admit(reqs[0])
    assert not tight.admit(reqs[2])  # 2500 particles and 3 bins do not fit beside it
    first = tight.drain()
    assert [r.req.req_id for r in first] == [reqs[0].req_id]
    _same(first[0], mixed[reqs[0].req_id], reqs[0].req_id)
    assert tight.device_bytes() <= lim + need * 1.2


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
        assert np.array_equal(res.footprint, mixed[rid].footprint)
        assert st.read_coord("release_lon")[0] == pytest.approx(res.req.lon)
        assert np.isfinite(res.latency_s)


def test_out_of_range_wave_slots_mark_unused_blocks(setup):
    """wave_slot entries outside [0, max_slots) (the footprint's n_r, the
    size of the slots array) mark unused blocks like -1: when a finished
    request's blocks below a running one's (so inside the launch) hold such
    values on the device, the running request's result is bit for bit the
    untouched server's."""
    cfg, met, ring, reqs = setup
    out = []
    for poke in (False, True):
        srv = _server(cfg, ring)
        assert srv.admit(reqs[1]) and srv.admit(reqs[0])  # blocks 0..10, then 11..34
        done = []
        while not any(r.req.req_id == reqs[1].req_id for r in done):
            srv.step()
            srv.poll()
            done += srv.take_completed()
        free = np.flatnonzero(srv.wave_slot_h[:srv.hi_blocks] < 0)
        assert free.size > 3  # the finished request's blocks, below the running one's
        if poke:
            bad = np.resize(np.array([srv.max_slots, srv.max_slots + 1000, 2**31 - 1], np.int32), free.size)
            dev = srv.wave_slot.device
            srv.wave_slot[torch.from_numpy(free).to(dev)] = torch.from_numpy(bad).to(dev)
            torch.cuda.synchronize()
            print(f"wave_slot: {free.size} free blocks set to {bad[:3].tolist()}", flush=True)
        (res,) = [r for r in done + srv.drain() if r.req.req_id == reqs[0].req_id]
        out.append(res)
    _same(out[0], out[1], "poke")
    assert out[0].total > 0
