"""CPU: batch schedules, release arrays and request streams (host logic of the
engine), and the per-release terminus pinned on the oracle."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from glide_amd.schedule import cursor_schedule, particle_ids, release_arrays
from glide_amd.synthetic import MetSpec, Release, SyntheticMet, serving_requests, workload_config
from oracle import run as ORUN


def _rel(t, length, n=10, idx=0, **kw):
    return Release(lon=5.0, lat=60.0, alt_agl_m=50.0, n_particles=n, release_time_s=t, duration_s=60,
                   length_s=length, seed=3 + idx, release_idx=idx, **kw)


def test_schedule_terminus_per_release_and_absolute_noise_step():
    a = _rel(10 * 3600 - 120, 3600, idx=0)
    b = _rel(12 * 3600 - 120, 5 * 3600, idx=1)  # ends later, reaches back further
    sched = cursor_schedule([a, b], 60)
    ends = [r.release_time_s + r.duration_s for r in (a, b)]
    term = min(e - r.length_s for e, r in zip(ends, (a, b)))
    assert sched[0].cursor_s == max(ends)
    assert sched[-1].cursor_s > term and sched[-1].cursor_s - sched[-1].delta_s <= term
    assert all(s.noise_step == s.cursor_s // 60 for s in sched)
    assert len({s.noise_step for s in sched}) == len(sched)
    # one release: the length override and the release's own length agree
    s1 = cursor_schedule([a], 60)
    s2 = cursor_schedule([dataclasses.replace(a, length_s=1)], 60, length_s=3600)
    assert [(x.cursor_s, x.delta_s, x.alpha) for x in s1] == [(x.cursor_s, x.delta_s, x.alpha) for x in s2]
    assert len(s1) == 60


def test_release_arrays_ids_and_offsets():
    rels = [_rel(3600, 7200, n=5, idx=0), _rel(7200, 3600, n=7, idx=1), _rel(5400, 3600, n=3, idx=2, noise_id=9)]
    ra = release_arrays(rels)
    assert ra.uid.size == 15 and np.unique(ra.uid).size == 15
    assert np.array_equal(ra.uid[:5], np.arange(5, dtype=np.uint32))
    assert np.array_equal(ra.uid[5:12], (np.uint32(1) << np.uint32(20)) + np.arange(7, dtype=np.uint32))
    assert np.array_equal(ra.uid[12:], (np.uint32(9) << np.uint32(20)) + np.arange(3, dtype=np.uint32))
    assert np.array_equal(particle_ids(rels[1]), ra.uid[5:12])
    length = np.repeat(np.float32([7200, 3600, 3600]), [5, 7, 3])
    assert np.array_equal(ra.lo_off, ra.r_off - length)
    off = np.repeat(np.float32([0, 3600, 1800]), [5, 7, 3])
    assert np.array_equal(ra.te_off, (off + np.float32(60)) - length)
    assert ra.lo_off.dtype == ra.te_off.dtype == np.float32
    assert np.all((ra.r_off >= off) & (ra.r_off < off + 60))


def test_network_and_serving_workloads_are_well_formed():
    cfg = workload_config("batched-gpu-336r")
    rels = cfg["releases"]
    assert len(rels) == 336 and cfg["grid"]["n_r"] == 336
    assert {r.length_s // 3600 for r in rels} <= {24, 48, 72, 96, 120}
    assert all(2000 <= r.n_particles <= 20000 and 10.0 <= r.alt_agl_m <= 300.0 for r in rels)
    assert min(r.release_time_s + r.duration_s - r.length_s for r in rels) >= 0  # met from hour 0
    assert max(r.release_time_s + r.duration_s for r in rels) // 3600 + 1 < cfg["spec"].n_hours
    spec = MetSpec(nx=240, ny=160, za=60, n_hours=25, lon0=2.0, lat0=72.0)
    reqs = serving_requests(spec, rate_per_s=20.0, seconds=5.0)
    arr = np.array([r.arrival_s for r in reqs])
    assert 60 <= len(reqs) <= 140 and np.all(np.diff(arr) > 0) and arr[-1] < 5.0
    for r in reqs:
        end = r.release_time_s + r.duration_s
        assert 6 * 3600 <= r.length_s <= 240 * 3600 and end - r.length_s >= 0 and end < 240 * 3600
        assert 1000 <= r.n_particles <= 50000 and 10.0 <= r.alt_agl_m <= 300.0
    assert len({r.req_id for r in reqs}) == len(reqs)


def test_oracle_release_terminus_equals_release_alone():
    """Per-release terminus: in a batch, a release whose terminus is not the
    batch's is stepped exactly as alone (the default mask steps it once more,
    main.py:1383)."""
    spec = MetSpec(nx=24, ny=20, za=30, n_hours=4, seed=77)
    met = SyntheticMet(spec)
    base = workload_config("plain-cpu-10k-3h")
    grid = dict(base["grid"], lon_min=float(spec.lon.min()), lon_max=float(spec.lon.max()) + 0.25,
                lat_min=float(spec.lat.min()), lat_max=float(spec.lat.max()) + 0.25, n_x=24, n_y=20, n_t=2, n_r=2)
    dom = dict(lon_min=float(spec.lon.min()), lon_max=float(spec.lon.max()), lat_min=float(spec.lat.min()),
               lat_max=float(spec.lat.max()), alt_max=20000.0)
    lon, lat = float(spec.lon[12]), float(spec.lat[10])
    a = Release(lon=lon, lat=lat, alt_agl_m=50.0, n_particles=40, release_time_s=3 * 3600 - 120, duration_s=60,
                length_s=3600, seed=5, release_idx=0)
    b = Release(lon=lon + 0.3, lat=lat, alt_agl_m=50.0, n_particles=30, release_time_s=2 * 3600 + 600,
                duration_s=60, length_s=5400, seed=6, release_idx=1)
    cfg = dict(base, spec=spec, releases=[a, b], grid=grid, domain=dom, convection=False, seed=21)
    both = ORUN.run(cfg, dty=np.float64, met=met, release_terminus=True)
    ref = ORUN.run(cfg, dty=np.float64, met=met)
    alone = ORUN.run(dict(cfg, releases=[a], grid=dict(grid, n_r=1)), dty=np.float64, met=met)
    sel = np.isin(both.particles["uid"], alone.particles["uid"])
    for k in ("lon", "lat", "z", "wp"):
        assert np.array_equal(both.particles[k][sel], alone.particles[k]), k
    # under the reference's batch terminus release a is stepped once more
    assert not np.array_equal(ref.particles["z"][sel], alone.particles["z"])
    assert sum(ref.active_per_step) > sum(both.active_per_step)


def test_met_ingest_rejects_bad_arguments():
    from glide_amd.ingest import MetIngest

    for kw in (dict(workers=0), dict(lookahead=-1), dict(max_resident=1)):
        with pytest.raises(ValueError, match="need workers"):
            MetIngest(object(), **kw)


def test_expand_to_batches_order_seeds_and_chunks():
    """RunConfig.expand_to_batches (config.py:530-621): time-major (time, site)
    order, seed + release_idx, chunks of max_releases_per_batch consecutive
    releases with global indices; a later batch's particles fill footprint
    slabs from 0 (main.py:1299-1302) and keep their global noise ids."""
    from glide_amd.schedule import Site, expand_to_batches

    sites = [Site(1.0, 50.0, 10.0, "a"), Site(2.0, 51.0, 20.0, "b"), Site(3.0, 52.0, 30.0, "c")]
    batches = expand_to_batches("multi_point_periodic", start_time_s=7200, period_s=3600, n_releases=4, sites=sites,
                                duration_s=60, n_particles=5, length_s=3600, seed=100, max_releases_per_batch=5)
    flat = [r for b in batches for r in b]
    assert [len(b) for b in batches] == [5, 5, 2]
    assert [r.release_idx for r in flat] == list(range(12))
    assert [r.seed for r in flat] == [100 + i for i in range(12)]
    assert [(r.release_time_s, r.lon) for r in flat] == [(7200 + 3600 * t, s.lon) for t in range(4) for s in sites]
    assert [r.alt_agl_m for r in flat[:3]] == [10.0, 20.0, 30.0]
    # per-release particle counts and lengths; no seed: every release seed 0 here (None in the reference)
    per = expand_to_batches("periodic_point", start_time_s=0, period_s=600, n_releases=3, lon=5.0, lat=6.0,
                            alt_agl_m=7.0, duration_s=60, n_particles=[3, 4, 5], length_s=[60, 120, 180], seed=None,
                            max_releases_per_batch=2)
    assert [[r.n_particles for r in b] for b in per] == [[3, 4], [5]]
    assert [r.length_s for b in per for r in b] == [60, 120, 180]
    assert [r.release_time_s for b in per for r in b] == [0, 600, 1200]
    sch = expand_to_batches("point_schedule", times_s=[50, 10, 30], lon=5.0, lat=6.0, alt_agl_m=7.0, duration_s=60,
                            n_particles=2, length_s=60, seed=1, max_releases_per_batch=10)
    assert [r.release_time_s for r in sch[0]] == [50, 10, 30] and [r.seed for r in sch[0]] == [1, 2, 3]
    (one,) = expand_to_batches("point", start_time_s=90, lon=5.0, lat=6.0, alt_agl_m=7.0, duration_s=60,
                               n_particles=8, length_s=60, seed=42, max_releases_per_batch=3)
    assert len(one) == 1 and one[0].seed == 42 and one[0].release_idx == 0
    with pytest.raises(TypeError):
        expand_to_batches("area", duration_s=60, n_particles=1, seed=0, max_releases_per_batch=1, length_s=60)
    # the third batch: slabs 0..1, noise ids 10..11, the same particle ids as in one batch of all 12
    ra = release_arrays(batches[2])
    assert np.array_equal(np.unique(ra.rel_idx), [0, 1])
    whole = release_arrays(flat)
    assert np.array_equal(ra.uid, whole.uid[-10:])
    assert np.array_equal(np.unique(whole.rel_idx), np.arange(12))


def test_network_releases_follow_expand_to_batches():
    """The synthetic measurement network is one multi-site periodic expansion:
    time-major, seeds seed + 1 + release_idx, one release per site per hour."""
    cfg = workload_config("batched-cpu-6r")
    rels = cfg["releases"]
    assert [r.release_idx for r in rels] == list(range(6))
    assert [r.seed for r in rels] == [rels[0].seed + i for i in range(6)]
    times = [r.release_time_s for r in rels]
    assert times == sorted(times) and len(set(times)) == 3  # 3 hours x 2 sites, time-major
    assert [(r.lon, r.lat) for r in rels[:2]] == [(r.lon, r.lat) for r in rels[2:4]]
