"""Batch schedule: release arrays and the cursor loop's per-step scalars.

Restates reference main.py:1338-1389 (cursor bounds and delta_s),
main.py:892-938 (bracket and alpha) and release_generator.py:240-322 (per-
release particle arrays). Times are integer seconds after the met epoch.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class StepInfo:
    index: int
    cursor_s: int  # pre-step cursor (s after the met epoch)
    delta_s: int
    cursor_off: float  # cursor - batch start
    bracket_hour: int  # hour_floor(cursor) as an hour index
    alpha: float
    noise_step: int  # Philox counter word 1: cursor_s // dt (absolute time, not the batch's step count)


def cursor_schedule(releases, dt: int, length_s: int | None = None) -> list[StepInfo]:
    """Steps of one batch: from the latest window end down to the earliest
    (window end - backward length) over the releases, ``while t_cursor >
    terminus``. ``length_s`` overrides every release's own length."""

    if dt <= 0:
        raise ValueError("dt_seconds must be > 0")
    ends = [r.release_time_s + r.duration_s for r in releases]
    lens = [int(length_s) if length_s is not None else int(r.length_s) for r in releases]
    start = max(ends)
    term = min(e - ln for e, ln in zip(ends, lens))
    batch_start = min(r.release_time_s for r in releases)
    out = []
    t = start
    i = 0
    while t > term:
        d = min(dt, t - term)
        bh = t // 3600
        alpha = max(0.0, min(1.0, ((t - 0.5 * d) - 3600.0 * bh) / 3600.0))
        out.append(StepInfo(i, int(t), int(d), float(t - batch_start), int(bh), alpha, int(t // dt)))
        t -= d
        i += 1
    return out


@dataclass
class ReleaseArrays:
    lon: np.ndarray
    lat: np.ndarray
    z: np.ndarray
    weight: np.ndarray
    r_off: np.ndarray
    we_off: np.ndarray
    rel_idx: np.ndarray
    uid: np.ndarray
    lo_off: np.ndarray  # r_off - backward length (float32): earliest active cursor offset
    te_off: np.ndarray  # window end - backward length (float32): the release's own terminus
    batch_start_s: int


def release_offsets(seed: int, n: int, duration_s: float) -> np.ndarray:
    """U[0, duration) float32 offsets from a CPU torch generator seeded with the
    release seed (release_generator.py:286-293)."""

    import torch

    g = torch.Generator(device="cpu")
    g.manual_seed(int(seed))
    return torch.empty(n, dtype=torch.float32).uniform_(0.0, float(duration_s), generator=g).numpy()


NOISE_ID_BITS = 12  # particle id = (noise_id % 4096) << 20 | k


def particle_ids(release) -> np.ndarray:
    """Philox counter word 0 of a release's particles: (noise id % 4096) << 20 |
    k for k < 2^20, so a release draws the same noise alone or in any batch of
    up to 4,096 releases."""

    n = int(release.n_particles)
    if n > (1 << 20):
        if n > (1 << 32) - 1:
            raise ValueError("a release holds at most 2^32 - 1 particles")
        return np.arange(n, dtype=np.uint32)  # one huge release: its own id space
    nid = release.release_idx if getattr(release, "noise_id", None) is None else release.noise_id
    return ((np.uint32(int(nid) % (1 << NOISE_ID_BITS)) << np.uint32(20)) + np.arange(n, dtype=np.uint32)).astype(np.uint32)


def release_arrays(releases) -> ReleaseArrays:
    """Particle arrays for a batch of point releases (weight 1/N per release,
    offsets relative to the earliest release in the batch). ``rel_idx`` is
    batch-local: a release's ``release_idx`` minus the first release's
    (main.py:1299-1302), so the batches of one expansion, whose indices are
    global, each fill footprint slabs from 0; particle ids keep the global
    index (the noise of a release does not depend on its batch)."""

    if not releases:
        raise ValueError("ReleaseBatch must contain at least one release")
    batch_start = min(r.release_time_s for r in releases)
    first_idx = int(releases[0].release_idx)
    cols = {k: [] for k in ("lon", "lat", "z", "weight", "r_off", "we_off", "rel_idx", "lo_off", "te_off")}
    for r in releases:
        n = int(r.n_particles)
        if n <= 0:
            raise ValueError(f"release {r.release_idx} has n_particles={n}; must be > 0")
        off = float(r.release_time_s - batch_start)
        cols["lon"].append(np.full(n, r.lon, np.float32))
        cols["lat"].append(np.full(n, r.lat, np.float32))
        cols["z"].append(np.full(n, r.alt_agl_m, np.float32))
        cols["weight"].append(np.full(n, 1.0 / n, np.float32))
        ro = release_offsets(r.seed, n, r.duration_s) + np.float32(off)
        cols["r_off"].append(ro)
        cols["lo_off"].append(ro - np.float32(r.length_s))
        cols["te_off"].append(np.full(n, np.float32(off + r.duration_s) - np.float32(r.length_s), np.float32))
        cols["we_off"].append(np.full(n, off + r.duration_s, np.float32))
        cols["rel_idx"].append(np.full(n, int(r.release_idx) - first_idx, np.int32))
    a = {k: np.concatenate(v) for k, v in cols.items()}
    uid = np.concatenate([particle_ids(r) for r in releases])
    return ReleaseArrays(uid=uid, batch_start_s=batch_start, **a)


@dataclass(frozen=True)
class Site:
    """One measurement site of a multi-site release (config.py: sites)."""

    lon: float
    lat: float
    alt_agl_m: float
    name: str = ""


def _per_release(v, idx: int):
    return v[idx] if isinstance(v, (list, tuple, np.ndarray)) else v


def expand_to_batches(kind: str, *, duration_s: int, n_particles, seed: int | None, max_releases_per_batch: int,
                      length_s, start_time_s: int = 0, period_s: int = 0, n_releases: int = 1, times_s=(),
                      lon: float | None = None, lat: float | None = None, alt_agl_m: float | None = None,
                      sites=()) -> list[list]:
    """The release schedule as execution batches (reference
    RunConfig.expand_to_batches, config.py:530-621).

    ``kind``: "point" (one release at ``start_time_s``), "periodic_point"
    (``n_releases`` every ``period_s`` from ``start_time_s``), "point_schedule"
    (one release per time in ``times_s``) at (``lon``, ``lat``,
    ``alt_agl_m``); "multi_point_periodic": one release per (time, site) in
    TIME-MAJOR order (every site of one release time, then the next time), so
    chunks of ``max_releases_per_batch`` consecutive releases span contiguous
    times across all sites. Release ``i`` (its index in that order) has seed
    ``seed + i`` (None when ``seed`` is None; a single point release keeps
    ``seed``). Releases are chunked into batches of at most
    ``max_releases_per_batch``. ``n_particles`` and ``length_s`` (the backward
    length; the reference has one per run) may be sequences indexed by
    release. Times are seconds after the met epoch. Returns a list of batches,
    each a list of ``glide_amd.synthetic.Release``."""

    from glide_amd.synthetic import Release

    if max_releases_per_batch < 1:
        raise ValueError("max_releases_per_batch must be >= 1")
    if kind == "point":
        return [[Release(lon=float(lon), lat=float(lat), alt_agl_m=float(alt_agl_m),
                         n_particles=int(_per_release(n_particles, 0)), release_time_s=int(start_time_s),
                         duration_s=int(duration_s), length_s=int(_per_release(length_s, 0)),
                         seed=0 if seed is None else int(seed), release_idx=0)]]
    if kind in ("periodic_point", "multi_point_periodic"):
        times = [int(start_time_s) + i * int(period_s) for i in range(int(n_releases))]
    elif kind == "point_schedule":
        times = [int(t) for t in times_s]
    else:
        raise TypeError(f"Unsupported release variant: {kind}")
    if kind == "multi_point_periodic":
        places = [(float(st.lon), float(st.lat), float(st.alt_agl_m)) for st in sites]
        if not places:
            raise ValueError("multi_point_periodic needs sites")
    else:
        places = [(float(lon), float(lat), float(alt_agl_m))]
    concretes = []
    idx = 0
    for t in times:  # time-major: every site of one release time, then the next time
        for plon, plat, palt in places:
            concretes.append(Release(lon=plon, lat=plat, alt_agl_m=palt,
                                     n_particles=int(_per_release(n_particles, idx)), release_time_s=t,
                                     duration_s=int(duration_s), length_s=int(_per_release(length_s, idx)),
                                     seed=(int(seed) + idx) if seed is not None else 0, release_idx=idx))
            idx += 1
    return [concretes[i:i + max_releases_per_batch] for i in range(0, len(concretes), max_releases_per_batch)]
