"""GPU: the serving workload at its spatial size. A bfloat16 ring of the
240 x 160 x 60 synthetic met (48 distinct hourly brackets, so requests here
reach back at most 40 h), per-column convection on, and a Poisson stream of
about 120 requests (20/s for 6 s; 1k-50k particles, 6-40 h back, 100 sites)
served in real time by the continuous-batching loop: every request completes
with a footprint, and for a sample of requests (the longest, the largest and
two more) the request served alone on a fresh server is bit for bit the one
served in the mix: particles, footprint, total, escape and active counts."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd import _lib as L  # noqa: E402
from glide_amd.runner import ConvectionOptions  # noqa: E402
from glide_amd.serving import MetRing, Server  # noqa: E402
from glide_amd.synthetic import MetSpec, SyntheticMet, serving_requests, workload_config  # noqa: E402

STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "alive")


def _by_uid(state):
    o = np.argsort(state["uid"].view(np.uint32), kind="stable")
    return {k: v[o] for k, v in state.items()}


def test_full_size_poisson_stream_alone_equals_mix():
    spec = MetSpec(nx=240, ny=160, za=60, n_hours=49, seed=20240105, lon0=2.0, lat0=72.0)
    met = SyntheticMet(spec)
    print("building the 48-bracket ring", flush=True)
    ring = MetRing(met, first_hour=0, n_hours=48, met_dtype=L.GLIDE_MET_BF16, convection=True)
    base = workload_config("plain-gpu-1M-24h")
    reqs = serving_requests(spec, rate_per_s=20.0, seconds=6.0, span_hours=48, length_h=(6.0, 40.0))
    assert len(reqs) >= 100

    def server():
        return Server(ring, base["grid"], base["domain"], max_slots=256, n_t_max=48, seed=29, keep_state=True,
                      convection=ConvectionOptions())

    srv = server()
    print(f"serving {len(reqs)} requests", flush=True)
    stats = srv.serve(reqs)
    res = {r.req.req_id: r for r in stats.results}
    assert sorted(res) == [r.req_id for r in reqs]
    assert all(r.footprint is not None and r.total > 0 and np.isfinite(r.latency_s) for r in res.values())
    assert srv.convect_launches > 0
    print(f"{stats.steps} steps, {stats.particle_steps:.3e} particle-steps in {stats.busy_s:.1f} s; "
          f"pool {srv.cap} particles, arena {srv.arena.size} cells", flush=True)
    longest = max(reqs, key=lambda r: r.length_s)
    largest = max(reqs, key=lambda r: r.n_particles)
    sample = {longest.req_id, largest.req_id, reqs[len(reqs) // 3].req_id, reqs[-1].req_id}
    for req in [r for r in reqs if r.req_id in sample]:
        alone_srv = server()
        assert alone_srv.admit(req)
        (alone,) = alone_srv.drain()
        mix = res[req.req_id]
        a, b = _by_uid(alone.state), _by_uid(mix.state)
        for k in STATE:
            assert np.array_equal(a[k], b[k]), (req.req_id, k)
        assert alone.escaped == mix.escaped and alone.active_steps == mix.active_steps
        assert np.array_equal(alone.footprint, mix.footprint) and alone.total == mix.total
        print(f"request {req.req_id} ({req.n_particles} particles, {req.length_s / 3600:.1f} h): alone == mix",
              flush=True)
