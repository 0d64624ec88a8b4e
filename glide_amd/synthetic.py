"""Seeded synthetic meteorology and release schedules for the step workloads.

The reference reads ERA5 from Zarr (met_reader.py); there is no network here, so
the workloads run on a generated met cube with the statistics the workload
definitions ask for. Everything is a deterministic function of the seed:

* grid: ``nx * ny`` columns at ``dlon``/``dlat`` degrees, latitude descending
  (ERA5 order), coordinates exact multiples of the spacing so the terrain
  gradient sees a uniform grid;
* source levels: 29 pressure levels 1000..50 hPa (top-of-atmosphere first, as
  ERA5 stores them), with per-column hydrostatic heights from a 6.5 K/km lapse
  rate, so some levels sit below the 0-500 m terrain and the resample must
  exclude them;
* u, v: 16 random 3-D Fourier modes, 10 m/s RMS, times a log-law boundary-layer
  profile; w ~ N(0, 0.05 m/s) Gaussian-smoothed; T with the 6.5 K/km lapse rate
  (isothermal above 11 km); q decaying with height;
* surface fields: blh on [300, 1500] m, u* on [0.1, 0.6] m/s, upward sensible
  heat flux on [-50, 200] W/m^2, surface pressure from the terrain;
* convection drivers per column: CAPE, LCL and cloud-top heights.

Snapshots differ by slow phase drift of every mode.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import cached_property

import numpy as np

GRAVITY = 9.80665
R_DRY = 287.05
LAPSE = 0.0065
P_SEA_PA = 101325.0
T_TROP_ALT_M = 11000.0

SOURCE_LEVELS_HPA = np.array(
    [50, 70, 100, 125, 150, 175, 200, 225, 250, 300, 350, 400, 450, 500, 550, 600, 650,
     700, 750, 775, 800, 825, 850, 875, 900, 925, 950, 975, 1000],
    dtype=np.float64,
)

CH3 = ("u", "v", "w", "t", "q")
CH2 = ("blh", "sp", "ustar", "shf")
CH_CONV = ("cape", "lcl_h", "top_h", "dtv")


def stretched_agl_levels(n_levels: int, alt_max_m: float, first_layer_m: float = 10.0) -> np.ndarray:
    """Geometric AGL ladder from 0 to ``alt_max_m`` with a ``first_layer_m``
    lowest layer (same construction as reference vertical_grid.py:70-144:
    200-step bisection on the stretch ratio, top pinned)."""

    if n_levels < 2:
        raise ValueError("n_levels must be >= 2")
    n_layers = n_levels - 1
    if n_layers == 1:
        return np.array([0.0, float(alt_max_m)])
    if first_layer_m * n_layers >= alt_max_m:
        raise ValueError("first_layer_m too thick for n_levels below alt_max_m")

    def span(r: float) -> float:
        if abs(r - 1.0) < 1e-12:
            return first_layer_m * n_layers
        return first_layer_m * (r**n_layers - 1.0) / (r - 1.0)

    lo, hi = 1.0, 2.0
    while span(hi) < alt_max_m:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if span(mid) < alt_max_m:
            lo = mid
        else:
            hi = mid
    r = 0.5 * (lo + hi)
    th = first_layer_m * r ** np.arange(n_layers, dtype=np.float64)
    lev = np.concatenate(([0.0], np.cumsum(th)))
    lev[-1] = float(alt_max_m)
    return lev


@dataclass(frozen=True)
class MetSpec:
    nx: int = 64
    ny: int = 48
    za: int = 40
    n_hours: int = 4
    seed: int = 1234
    lon0: float = 2.0
    lat0: float = 62.0  # northernmost row (latitude descends with the row index)
    dlon: float = 0.25
    dlat: float = -0.25
    alt_max_m: float = 20000.0
    first_layer_m: float = 10.0
    n_modes: int = 16
    wind_rms: float = 10.0
    # source_units: hours in the reader's raw units (omega in Pa/s for w,
    # geopotential and surface geopotential instead of heights, sensible heat
    # flux accumulated over the hour in J/m^2, downward-positive), for the
    # device conversions (glide_source_units)
    source_units: bool = False

    @property
    def lon(self) -> np.ndarray:
        return self.lon0 + self.dlon * np.arange(self.nx, dtype=np.float64)

    @property
    def lat(self) -> np.ndarray:
        return self.lat0 + self.dlat * np.arange(self.ny, dtype=np.float64)

    @property
    def agl_levels(self) -> np.ndarray:
        return stretched_agl_levels(self.za, self.alt_max_m, self.first_layer_m)


@dataclass
class SourceHour:
    """One hourly snapshot on the source (pressure) levels, as the reader hands
    it to the resample: 3-D channels [C3, Zp, Y, X] (u, v, w, t, q), 2-D channels
    [C2, Y, X] (blh, sp, ustar, shf), per-column AGL heights [Zp, Y, X] (not
    clamped: negative below ground) and the level pressures [Zp] in Pa."""

    src3: np.ndarray
    src2: np.ndarray
    height_agl: np.ndarray | None
    p_level_pa: np.ndarray
    conv: np.ndarray  # [4, Y, X] cape, lcl_h, top_h, dtv
    # "source": the reader's raw units (w channel omega in Pa/s, shf channel
    # J/m^2 accumulated over 3600 s and downward-positive, geopotential fields
    # in m^2/s^2 and height_agl None); "final": the units above
    units: str = "final"
    geopotential: np.ndarray | None = None  # [Zp, Y, X]
    geopotential_sfc: np.ndarray | None = None  # [Y, X]


def _smooth_pattern(rng, ny, nx, n=4, kmax=3.0):
    """Random smooth field generator: returns (kx, ky, phase, omega) for ``n`` modes."""

    kx = rng.uniform(0.5, kmax, n) * 2 * np.pi / nx * rng.choice([-1, 1], n)
    ky = rng.uniform(0.5, kmax, n) * 2 * np.pi / ny * rng.choice([-1, 1], n)
    ph = rng.uniform(0, 2 * np.pi, n)
    om = rng.uniform(0.2, 1.0, n) * 2 * np.pi / (48 * 3600.0)
    return kx, ky, ph, om


def _eval_pattern(p, xx, yy, t):
    kx, ky, ph, om = p
    s = np.zeros(xx.shape, dtype=np.float64)
    for m in range(kx.size):
        s += np.sin(kx[m] * xx + ky[m] * yy + ph[m] + om[m] * t)
    return np.tanh(s / np.sqrt(kx.size / 2.0))  # roughly uniform on (-1, 1)


class SyntheticMet:
    """Deterministic synthetic met source for a :class:`MetSpec`."""

    def __init__(self, spec: MetSpec):
        self.spec = spec
        rng = np.random.default_rng(np.random.PCG64(spec.seed))
        ny, nx = spec.ny, spec.nx
        self._yy, self._xx = np.meshgrid(np.arange(ny, dtype=np.float64),
                                         np.arange(nx, dtype=np.float64), indexing="ij")
        # Terrain: smooth, 0-500 m.
        tp = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx, n=5, kmax=2.0)
        terr = _eval_pattern(tp, self._xx, self._yy, 0.0)
        self.terrain = (250.0 + 250.0 * terr).astype(np.float32)  # [Y, X]
        self._tsfc = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx)
        self._blh = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx)
        self._ust = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx)
        self._shf = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx)
        self._qs = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx)
        self._cape = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx)
        self._lcl = _smooth_pattern(np.random.default_rng(rng.integers(1 << 62)), ny, nx)
        # Wind modes: 16 per component, horizontal wavelengths ~ 8..60 cells,
        # vertical wavelengths 6-30 km, periods 1-4 days.
        wr = np.random.default_rng(rng.integers(1 << 62))
        M = spec.n_modes
        self._wm = []
        for _ in range(2):
            kx = wr.uniform(0.1, 0.8, M) * wr.choice([-1, 1], M)
            ky = wr.uniform(0.1, 0.8, M) * wr.choice([-1, 1], M)
            kz = 2 * np.pi / wr.uniform(6000.0, 30000.0, M)
            ph = wr.uniform(0, 2 * np.pi, M)
            om = 2 * np.pi / (wr.uniform(1.0, 4.0, M) * 86400.0)
            amp = wr.uniform(0.5, 1.5, M)
            amp *= spec.wind_rms / np.sqrt(0.5 * np.sum(amp**2))
            self._wm.append((kx, ky, kz, ph, om, amp))
        # w: two smooth random fields per source level, rotated in time.
        from scipy.ndimage import gaussian_filter

        Zp = SOURCE_LEVELS_HPA.size
        wa = gaussian_filter(wr.standard_normal((Zp, ny, nx)), sigma=(0, 2, 2))
        wb = gaussian_filter(wr.standard_normal((Zp, ny, nx)), sigma=(0, 2, 2))
        wa *= 0.05 / wa.std()
        wb *= 0.05 / wb.std()
        self._wa = wa
        self._wb = wb
        self._w_omega = 2 * np.pi / (36 * 3600.0)

    @cached_property
    def lon(self) -> np.ndarray:
        return self.spec.lon

    @cached_property
    def lat(self) -> np.ndarray:
        return self.spec.lat

    def _uniform(self, pat, lo, hi, t):
        return lo + (hi - lo) * 0.5 * (1.0 + _eval_pattern(pat, self._xx, self._yy, t))

    def hour(self, h: int) -> SourceHour:
        """Source-level snapshot at hour index ``h`` (t = 3600 h seconds)."""

        s = self.spec
        t = 3600.0 * float(h)
        ny, nx = s.ny, s.nx
        p_hpa = SOURCE_LEVELS_HPA
        Zp = p_hpa.size
        tsfc = 288.0 + 8.0 * _eval_pattern(self._tsfc, self._xx, self._yy, t)  # sea-level T
        t_trop = tsfc - LAPSE * T_TROP_ALT_M
        expo = R_DRY * LAPSE / GRAVITY
        p_trop = P_SEA_PA * (t_trop / tsfc) ** (1.0 / expo)
        p = p_hpa[:, None, None] * 100.0
        z_lapse = (tsfc / LAPSE) * (1.0 - (p / P_SEA_PA) ** expo)
        z_strat = T_TROP_ALT_M + (R_DRY * t_trop / GRAVITY) * np.log(p_trop / p)
        z_msl = np.where(p >= p_trop, z_lapse, z_strat)
        temp = np.where(z_msl < T_TROP_ALT_M, tsfc - LAPSE * z_msl, t_trop)
        terrain = self.terrain.astype(np.float64)
        agl = (z_msl - terrain[None]).astype(np.float32)
        zbar = z_msl.mean(axis=(1, 2))
        # Log-law boundary-layer profile (z0 = 0.1 m, saturating at 1 km AGL).
        z0 = 0.1
        prof = np.clip(np.log1p(np.maximum(agl, 0.0) / z0) / np.log1p(1000.0 / z0), 0.0, 1.0)
        winds = []
        for kx, ky, kz, ph, om, amp in self._wm:
            th_h = (kx[:, None, None] * self._xx[None] + ky[:, None, None] * self._yy[None]
                    + (ph + om * t)[:, None, None])  # [M, Y, X]
            th_z = kz[None, :] * zbar[:, None]  # [Zp, M]
            A = np.concatenate([amp[None] * np.cos(th_z), amp[None] * np.sin(th_z)], axis=1)
            B = np.concatenate([np.sin(th_h), np.cos(th_h)], axis=0).reshape(2 * kx.size, -1)
            winds.append((A @ B).reshape(Zp, ny, nx) * prof)
        c, sn = np.cos(self._w_omega * t), np.sin(self._w_omega * t)
        w = c * self._wa + sn * self._wb
        qs = self._uniform(self._qs, 0.004, 0.016, t)
        q = qs[None] * np.exp(-np.maximum(agl, 0.0) / 2500.0)
        src3 = np.stack([winds[0], winds[1], w, temp, q]).astype(np.float32)
        blh = self._uniform(self._blh, 300.0, 1500.0, t)
        ust = self._uniform(self._ust, 0.1, 0.6, t)
        shf = self._uniform(self._shf, -50.0, 200.0, t)
        sp = P_SEA_PA * (1.0 - LAPSE * terrain / tsfc) ** (GRAVITY / (R_DRY * LAPSE))
        src2 = np.stack([blh, sp, ust, shf]).astype(np.float32)
        cape = np.maximum(self._uniform(self._cape, -1500.0, 3000.0, t), 0.0)
        lcl = self._uniform(self._lcl, 500.0, 2000.0, t)
        top = np.minimum(lcl + 1000.0 + 3.0 * cape, 15000.0)
        dtv = cape / 300.0
        conv = np.stack([cape, lcl, top, dtv]).astype(np.float32)
        if not s.source_units:
            return SourceHour(src3=src3, src2=src2, height_agl=agl, p_level_pa=p_hpa * 100.0, conv=conv)
        # the same hour in the reader's raw units (inverse conversions, float64, to float32)
        omega = -(src3[2].astype(np.float64) * GRAVITY * p / (R_DRY * src3[3].astype(np.float64)))
        src3[2] = omega.astype(np.float32)
        src2[3] = (-(src2[3].astype(np.float64) * 3600.0)).astype(np.float32)
        return SourceHour(src3=src3, src2=src2, height_agl=None, p_level_pa=p_hpa * 100.0, conv=conv,
                          units="source", geopotential=(z_msl * GRAVITY).astype(np.float32),
                          geopotential_sfc=(terrain * GRAVITY).astype(np.float32))


@dataclass(frozen=True)
class Release:
    """One point release (reference ConcreteRelease, config.py:225-240)."""

    lon: float
    lat: float
    alt_agl_m: float
    n_particles: int
    release_time_s: int  # seconds after the met epoch (hour 0)
    duration_s: int = 60
    length_s: int = 3 * 3600  # backward integration length
    seed: int = 0
    release_idx: int = 0  # index in the run's release order; minus the batch's first: its footprint slab
    noise_id: int | None = None  # particle ids are (noise_id % 4096) << 20 | k; None = release_idx


def _domain_and_grid(spec: MetSpec, n_r: int, n_t: int):
    lon_lo, lon_hi = float(spec.lon.min()), float(spec.lon.max())
    lat_lo, lat_hi = float(spec.lat.min()), float(spec.lat.max())
    grid = dict(
        lon_min=lon_lo, lon_max=lon_hi + abs(spec.dlon), lat_min=lat_lo,
        lat_max=lat_hi + abs(spec.dlat), n_x=spec.nx, n_y=spec.ny, n_t=n_t, n_r=n_r,
        z_edges=[0.0, 100.0], n_z=1,
    )
    domain = dict(lon_min=lon_lo, lon_max=lon_hi, lat_min=lat_lo, lat_max=lat_hi,
                  alt_max=spec.alt_max_m)
    return grid, domain


def network_releases(spec: MetSpec, *, n_sites: int, n_hourly: int, first_hour: int, lengths_h,
                     n_range: tuple[int, int], seed: int, duration_s: int = 60) -> list[Release]:
    """A measurement network's hourly releases as one batch: ``n_sites`` sites
    at seeded interior positions with inlet heights uniform on [10, 300] m,
    one release per site per hour for ``n_hourly`` hours from ``first_hour``
    (2 minutes before the hour), in the reference's time-major order with
    seeds ``seed + release_idx`` (``schedule.expand_to_batches``,
    "multi_point_periodic"); per release a particle count log-uniform on
    ``n_range`` and a backward length drawn from ``lengths_h`` (hours)."""

    from glide_amd.schedule import Site, expand_to_batches

    rng = np.random.default_rng(seed)
    lon, lat = spec.lon, spec.lat
    sites = []
    for _ in range(n_sites):
        slon = float(rng.uniform(lon[len(lon) // 4], lon[3 * len(lon) // 4]))
        slat = float(rng.uniform(min(lat[len(lat) // 4], lat[3 * len(lat) // 4]),
                                 max(lat[len(lat) // 4], lat[3 * len(lat) // 4])))
        sites.append(Site(lon=slon, lat=slat, alt_agl_m=float(rng.uniform(10.0, 300.0))))
    n_rel = n_sites * n_hourly
    counts = [int(round(float(np.exp(v)))) for v in rng.uniform(np.log(n_range[0]), np.log(n_range[1]), n_rel)]
    lengths = [int(v) * 3600 for v in rng.choice(np.asarray(lengths_h), n_rel)]
    (batch,) = expand_to_batches("multi_point_periodic", start_time_s=first_hour * 3600 - 120, period_s=3600,
                                 n_releases=n_hourly, sites=sites, duration_s=duration_s, n_particles=counts,
                                 length_s=lengths, seed=seed + 1, max_releases_per_batch=n_rel)
    return batch


def workload_config(name: str):
    """(MetSpec, [Release], grid dict, n_substeps, convection flag) for the named
    workloads:

    * ``plain-cpu-10k-3h``: 10,000 particles at 50 m, 3 h back, dt = 60 s, 64x48
      grid at 0.25 deg, 40 AGL levels to 20 km, 4 hourly snapshots, convection
      off, footprint 64x48 x one 0-100 m layer x 3 hourly time-ago bins.
    * ``plain-gpu-1M-24h``: 2^20 particles, 24 h back, 240x160 grid, 60 levels,
      25 snapshots, convection on, footprint 24 time-ago bins x one layer.
    * ``batched-gpu-336r``: one GPU's share (336 of 2,688) of the network batch:
      7 sites x 48 hourly releases, 2k-20k particles each (log-uniform), 24 to
      120 h back, 240x160x60 met over 169 hourly snapshots, convection on,
      footprint 336 releases x 120 time-ago bins x one layer.
    * ``batched-gpu-2688r`` (float32 met) / ``batched-gpu-2688r-bf16`` (the
      same releases; the precision layout is chosen by the engine): the full
      network batch, 56 sites x 48 hourly releases (2,688), 2k-20k particles
      each, 24 to 120 h back, 240x160x60 met over 169 hourly snapshots,
      convection on, footprint 2,688 releases x up to 120 time-ago bins x one
      layer in per-release slabs.
    * ``batched-cpu-6r``: a small network batch on the 64x48x40 grid: 2 sites x
      3 hourly releases of 8k-16k particles, 1 to 3 h back, convection off.
    """

    if name in ("plain-cpu-10k-3h", "plain-gpu-1M-24h"):
        if name == "plain-cpu-10k-3h":
            spec = MetSpec(nx=64, ny=48, za=40, n_hours=4, seed=20240101)
            n, hours, n_t, conv = 10_000, 3, 3, False
        else:
            spec = MetSpec(nx=240, ny=160, za=60, n_hours=25, seed=20240102, lon0=2.0, lat0=72.0)
            n, hours, n_t, conv = 1 << 20, 24, 24, True
        lon_c = float(spec.lon[(spec.nx - 1) // 2])
        lat_c = float(spec.lat[(spec.ny - 1) // 2])
        length = hours * 3600
        rel = Release(lon=lon_c + 0.1, lat=lat_c + 0.1, alt_agl_m=50.0, n_particles=n,
                      release_time_s=length - 120, duration_s=60, length_s=length, seed=7)
        grid, domain = _domain_and_grid(spec, 1, n_t)
        return dict(spec=spec, releases=[rel], grid=grid, domain=domain, convection=conv,
                    dt=60, seed=rel.seed)
    if name == "batched-gpu-336r":
        spec = MetSpec(nx=240, ny=160, za=60, n_hours=169, seed=20240103, lon0=2.0, lat0=72.0)
        rels = network_releases(spec, n_sites=7, n_hourly=48, first_hour=121, lengths_h=(24, 48, 72, 96, 120),
                                n_range=(2000, 20000), seed=20240103)
        grid, domain = _domain_and_grid(spec, len(rels), 120)
        return dict(spec=spec, releases=rels, grid=grid, domain=domain, convection=True, dt=60, seed=11)
    if name in ("batched-gpu-2688r", "batched-gpu-2688r-bf16"):
        # The network batch at its full size: 56 sites x 48 hourly releases, one
        # engine; 168 hourly snapshots cover the 48 release hours and the
        # longest (120 h) backward window.
        spec = MetSpec(nx=240, ny=160, za=60, n_hours=169, seed=20240106, lon0=2.0, lat0=72.0)
        rels = network_releases(spec, n_sites=56, n_hourly=48, first_hour=121, lengths_h=(24, 48, 72, 96, 120),
                                n_range=(2000, 20000), seed=20240106)
        grid, domain = _domain_and_grid(spec, len(rels), 120)
        return dict(spec=spec, releases=rels, grid=grid, domain=domain, convection=True, dt=60, seed=23)
    if name == "batched-cpu-6r":
        spec = MetSpec(nx=64, ny=48, za=40, n_hours=7, seed=20240104)
        rels = network_releases(spec, n_sites=2, n_hourly=3, first_hour=4, lengths_h=(1, 2, 3),
                                n_range=(8000, 16000), seed=20240104)
        grid, domain = _domain_and_grid(spec, len(rels), 3)
        return dict(spec=spec, releases=rels, grid=grid, domain=domain, convection=False, dt=60, seed=13)
    raise KeyError(f"unknown workload {name!r}")


def serving_requests(spec: MetSpec, *, rate_per_s: float, seconds: float, n_sites: int = 100,
                     span_hours: int = 240, length_h=(6.0, 240.0), n_range=(1000, 50000), seed: int = 20240105):
    """A Poisson stream of footprint requests (the serving workload): arrivals
    at ``rate_per_s`` over ``seconds``; each request picks one of ``n_sites``
    seeded sites, an inlet height uniform on [10, 300] m, a backward length
    uniform on ``length_h`` (whole minutes), a release time uniform within the
    ``span_hours`` met span that leaves its backward window inside the span,
    and a particle count log-uniform on ``n_range``."""

    from glide_amd.serving import Request

    rng = np.random.default_rng(seed)
    lon, lat = spec.lon, spec.lat
    s_lon = rng.uniform(lon[len(lon) // 8], lon[7 * len(lon) // 8], n_sites)
    lat_lo, lat_hi = sorted((lat[len(lat) // 8], lat[7 * len(lat) // 8]))
    s_lat = rng.uniform(lat_lo, lat_hi, n_sites)
    out = []
    t = 0.0
    k = 0
    span = span_hours * 3600
    while True:
        t += float(rng.exponential(1.0 / rate_per_s))
        if t >= seconds:
            break
        site = int(rng.integers(n_sites))
        length = int(round(float(rng.uniform(*length_h)) * 60.0)) * 60
        length = min(length, span - 3 * 3600)
        rel = int(rng.integers(length + 60, span - 120))
        n = int(round(float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))))
        out.append(Request(req_id=k, lon=float(s_lon[site]), lat=float(s_lat[site]),
                           alt_agl_m=float(rng.uniform(10.0, 300.0)), n_particles=n, release_time_s=rel,
                           length_s=length, duration_s=60, seed=seed + 1 + k, arrival_s=t))
        k += 1
    return out
