"""The plain FP32 method as PyTorch ops: the comparison baseline.

This is the reference's own step restated with the same torch operations it
uses (HannaScheme static path, hanna.py:855-1068, and its CPU dynamic path,
hanna.py:635-745; GPUEngine.normalize_particle_coordinates,
gpu_engine.py:210-299; EmanuelReducedConvection, emanuel.py:570-805, applied
per column; the driver's mask, gridding and escape kill, main.py:1396-1477,
footprint_gridder.py:117-197): trilinear and bilinear
``grid_sample(align_corners=True)``, a ``searchsorted`` level lookup,
``torch.randn`` noise, the substep loop, the materialised transposed mass-flux
rows with cumulative-sum sampling, and ``scatter_add_`` gridding with float
atomics. Every operation inside a step is a PyTorch kernel; none of this
build's HIP kernels run inside a step. The hourly window (winds of both
bracket hours, surface fields, the support stacks, the convection inputs) is
handed in per met hour by a window provider, as the reference builds its
static buffers once per window.

Three modes, the reference's three ways of running the same step:
* static (default): every particle, the fixed ``max_substeps`` loop with
  finished particles advanced by ``sub_dt = 0`` (the GPU static path);
* ``dynamic=True``: the active subset only (boolean-indexed) and a loop of
  max(k) substeps (the reference's default CPU path);
* ``graph=True`` (GPU): the static step body captured once as a HIP graph
  over address-stable buffers, the per-step scalars (cursor, dt, alpha) in
  device tensors refreshed between replays (the reference's
  ``_maybe_graph_compile`` replay, hanna.py:1117-1160); the first step (which
  initialises every op before capture) and, with convection on, the first step
  of each met hour run eagerly.

Noise comes from the device's default torch generator, seeded by ``reset``
(so it reseeds torch's global RNG for that device); a replay draws the same
numbers as the eager body would at that point of the sequence.

It exists to measure the speedup of the HIP engine over the plain method and
to time the reference's arithmetic on the host's CPU cores; it is not a
product path and not a correctness oracle (its noise is torch's).
"""

from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib as L

KARMAN, GRAVITY, R_DRY, C_P, OMEGA = 0.4, 9.80665, 287.05, 1005.0, 7.2921e-5
SIGMA_MIN, T_L_MIN, USTAR_MIN, BLH_MIN = 1e-3, 1.0, 1e-3, 50.0
M_LAT, M_LON_EQ = 110540.0, 111320.0
CFL_MAX_OUTFLUX_FRAC = 0.9

WINDOW_FIELDS = ("uvw", "ft", "dens", "mea", "surf_s", "surf_e", "t_s", "t_e", "conv", "p_hpa")


class TorchPlain:
    """One release batch stepped with PyTorch ops on ``device``.

    ``windows(bracket_hour)`` returns the hour's fields as tensors on the
    device: ``uvw`` [6, Z, Y, X] (u v w at the start, then at the end of the
    bracket), ``ft`` / ``dens`` / ``mea`` [2, Z, Y, X] (free-troposphere sigma
    and T_L, density and its vertical gradient, meander sigmas), ``surf_s`` /
    ``surf_e`` [5, Y, X] (blh sp ustar shf t0), ``t_s`` / ``t_e`` [Z, Y, X]
    (temperature), ``conv`` [4, Y, X] (CAPE, LCL and cloud-top heights, T_v
    excess) and ``p_hpa`` [Z]; ``geometry`` gives the grid's first and last
    lon / lat."""

    def __init__(self, cfg: dict, windows, *, device, levels, geometry: dict, seed: int = 0, hanna=None,
                 convection: bool = True, dynamic: bool = False, graph: bool = False):
        from .runner import HannaOptions
        from .schedule import cursor_schedule, release_arrays

        self.dev = torch.device(device)
        if graph and self.dev.type != "cuda":
            raise ValueError("graph capture needs a GPU device")
        if graph and dynamic:
            raise ValueError("the captured step body is the static path")
        self.cfg = cfg
        self.windows = windows
        self.h = hanna or HannaOptions()
        self.convection = bool(convection) and bool(cfg.get("convection", True))
        self.dynamic = bool(dynamic)
        self.graph = bool(graph)
        self.sched = cursor_schedule(cfg["releases"], int(cfg["dt"]))
        ra = release_arrays(cfg["releases"])
        d = self.dev
        t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(d, dt)  # noqa: E731
        self._init = dict(lon=t(ra.lon), lat=t(ra.lat), z=t(ra.z))
        self.weight, self.r_off, self.we_off = t(ra.weight), t(ra.r_off), t(ra.we_off)
        self.lo_off = t(ra.lo_off)
        self.rel = t(ra.rel_idx, torch.int64)
        n = int(ra.lon.size)
        self.n = n
        g = cfg["grid"]
        self.fg = g
        self.dom = cfg["domain"]
        self.z_edges = torch.tensor(g["z_edges"], dtype=torch.float32, device=d)
        self.n_cells = g["n_r"] * g["n_t"] * g["n_z"] * g["n_y"] * g["n_x"]
        self.levels = torch.as_tensor(np.asarray(levels, np.float32), device=d)
        self.lon0, self.lon1 = float(geometry["lon0"]), float(geometry["lon1"])
        self.lat0, self.lat1 = float(geometry["lat0"]), float(geometry["lat1"])
        # address-stable state (a captured body reads and writes these in place)
        self.lon, self.lat, self.z = (torch.empty(n, device=d) for _ in range(3))
        self.up, self.vp, self.wp, self.um, self.vm = (torch.empty(n, device=d) for _ in range(5))
        self.alive = torch.empty(n, dtype=torch.bool, device=d)
        self.footprint = torch.zeros(self.n_cells, device=d)
        # per-particle step counts (elementwise adds: a captured body holds no reduction)
        self.act_count = torch.zeros(n, dtype=torch.int32, device=d)
        # per-step scalars as 0-d tensors (refreshed in place between replays)
        self.co = torch.zeros((), device=d)
        self.dt = torch.zeros((), device=d)
        self.al = torch.zeros((), device=d)
        self.win: dict[str, torch.Tensor] | None = None
        self.seed = int(seed)
        self._graph = None
        self._warm = False
        self.reset()

    # -- construction from an engine's met (GPU) ---------------------------
    @classmethod
    def from_engine(cls, eng, *, seed: int = 0, convection: bool = True, graph: bool = False,
                    dynamic: bool = False) -> "TorchPlain":
        """The engine's workload, with each hour's window taken from its
        resident met (the packed window it builds, unpacked to fields)."""

        if eng.met.met_dtype != L.GLIDE_MET_F32:
            raise ValueError("the plain method reads float32 met")
        m = eng.met

        def windows(bh: int) -> dict:
            st = L.stream_ptr()
            m.build_window(bh, st)
            nz, ny, nx = m.nz, m.ny, m.nx
            wind, aux = L.unpack_window(m.wind, m.aux, nz, ny, nx, m.met_dtype)
            surf = m.surf.view(torch.float32).view(ny, nx, 12)
            s3, _, p_hpa = m.resampled(bh, st)
            e3, _, _ = m.resampled(bh + 1, st)
            f = lambda a: a.permute(3, 0, 1, 2).contiguous()  # noqa: E731  [Z,Y,X,C] -> [C,Z,Y,X]
            return dict(uvw=f(wind), ft=f(aux[..., 0:2]), dens=f(aux[..., 2:4]), mea=f(aux[..., 4:6]),
                        surf_s=surf[..., 0:5].permute(2, 0, 1).contiguous(),
                        surf_e=surf[..., 5:10].permute(2, 0, 1).contiguous(),
                        t_s=s3[3].contiguous(), t_e=e3[3].contiguous(), conv=m.source(bh, st)["conv"],
                        p_hpa=p_hpa.to(torch.float64))

        g = m.grid
        return cls(eng.cfg, windows, device=eng.device, levels=m.levels_np,
                   geometry=dict(lon0=g.lon0, lon1=g.lon1, lat0=g.lat0, lat1=g.lat1), seed=seed, hanna=eng.h,
                   convection=convection and eng.conv_opts is not None, graph=graph, dynamic=dynamic)

    # -- state ------------------------------------------------------------
    def reset(self) -> None:
        """Back to the release state (a captured body stays valid)."""

        for k in ("lon", "lat", "z"):
            getattr(self, k).copy_(self._init[k])
        for k in ("up", "vp", "wp", "um", "vm"):
            getattr(self, k).zero_()
        self.alive.fill_(True)
        self.footprint.zero_()
        self.act_count.zero_()
        # the device's default generator, as the reference draws its noise;
        # a captured body advances it exactly as the eager body does
        if self.dev.type == "cuda":
            torch.cuda.manual_seed(self.seed)
        else:
            torch.manual_seed(self.seed)
        self.step_i = 0
        self._win_hour = None

    def _load_window(self, bh: int) -> None:
        fields = self.windows(bh)
        if self.win is None:
            self.win = {k: fields[k].to(self.dev).clone() for k in WINDOW_FIELDS}
        else:  # in place: a captured body keeps reading the same buffers
            for k in WINDOW_FIELDS:
                self.win[k].copy_(fields[k])
        self._win_hour = bh

    # -- interpolation (gpu_engine.py:210-299, hanna.py:1070-1085) --------
    def _norm(self, lon, lat, z):
        n = self.levels.numel()
        up = torch.searchsorted(self.levels, z.contiguous()).clamp(1, n - 1)
        lo = up - 1
        l_lo, l_hi = self.levels[lo], self.levels[up]
        frac = ((z - l_lo) / torch.clamp(l_hi - l_lo, min=1e-6)).clamp(0.0, 1.0)
        zn = 2.0 * (lo.to(z.dtype) + frac) / float(n - 1) - 1.0
        xn = 2.0 * ((lon - self.lon0) / (self.lon1 - self.lon0)) - 1.0
        yn = 2.0 * ((lat - self.lat0) / (self.lat1 - self.lat0)) - 1.0
        return torch.stack([xn.clamp(-1, 1), yn.clamp(-1, 1), zn.clamp(-1, 1)], -1)

    @staticmethod
    def _sample3(field, grid):
        out = F.grid_sample(field.unsqueeze(0), grid.view(1, -1, 1, 1, 3), mode="bilinear", padding_mode="zeros",
                            align_corners=True)
        return out.view(field.shape[0], -1)

    def _sample2(self, field, lon, lat):
        xn = (2.0 * (lon - self.lon0) / (self.lon1 - self.lon0) - 1.0).clamp(-1, 1)
        yn = (2.0 * (lat - self.lat0) / (self.lat1 - self.lat0) - 1.0).clamp(-1, 1)
        g = torch.stack([xn, yn], -1).view(1, -1, 1, 2)
        out = F.grid_sample(field.unsqueeze(0), g, mode="bilinear", padding_mode="zeros", align_corners=True)
        return out.view(field.shape[0], -1)

    # -- Hanna closure (hanna.py:142-426, 1500-1596) -----------------------
    def _column(self, z, blh, ustar, wstar, hol, f, ft_grid_z):
        zoh = (z / blh).clamp(0.0, 1.0)
        # stable
        om = torch.clamp(1.0 - zoh, min=0.0)
        su_s, sw_s = 2.0 * ustar * om, 1.3 * ustar * om
        zf = torch.clamp(zoh, min=1e-6)
        tu_s = 0.15 * blh / torch.clamp(su_s, min=SIGMA_MIN) * torch.sqrt(zf)
        tv_s = 0.07 * blh / torch.clamp(sw_s, min=SIGMA_MIN) * torch.sqrt(zf)
        tw_s = 0.10 * blh / torch.clamp(sw_s, min=SIGMA_MIN) * zf.pow(0.8)
        # neutral
        us = torch.clamp(ustar, min=USTAR_MIN)
        fa = torch.clamp(f.abs(), min=1e-5)
        su_n = 2.0 * ustar * torch.exp(-3.0 * fa * z / us)
        sw_n = 1.3 * ustar * torch.exp(-2.0 * fa * z / us)
        tw_n = 0.5 * z / torch.clamp(sw_n, min=SIGMA_MIN) / torch.clamp(1.0 + 15.0 * fa * z / us, min=1.0)
        # unstable
        hc = hol.clamp(-1000.0, 0.0)
        su_u = torch.pow(torch.clamp(12.0 - 0.5 * hc, min=1e-3), 1.0 / 3.0) * ustar
        sw2 = 1.2 * wstar ** 2 * torch.clamp(1.0 - 0.9 * zoh, min=0.0) * zoh.pow(2.0 / 3.0) + (1.8 - 1.4 * zoh) * ustar ** 2
        sw_u = torch.sqrt(torch.clamp(sw2, min=SIGMA_MIN ** 2))
        tu_u = 0.15 * blh / torch.clamp(su_u, min=SIGMA_MIN)
        azl = torch.clamp(zoh * (-hc), min=0.0)
        swc = torch.clamp(sw_u, min=SIGMA_MIN)
        tw_u = torch.where(azl < 1.0, 0.1 * z / (swc * torch.clamp(0.55 - 0.38 * azl, min=1e-2)),
                           torch.where(zoh < 0.1, 0.59 * z / swc, 0.15 * blh / swc * (1.0 - torch.exp(-5.0 * zoh))))
        st, un = hol > 1.0, hol < -1.0
        pick = lambda a, b, c: torch.where(st, a, torch.where(un, c, b))  # noqa: E731
        su = pick(su_s, su_n, su_u)
        sv = pick(sw_s, sw_n, su_u)
        sw = pick(sw_s, sw_n, sw_u)
        tu = pick(tu_s, tw_n, tu_u)
        tv = pick(tv_s, tw_n, tu_u)
        tw = pick(tw_s, tw_n, tw_u)
        if self.h.flexpart_tl_floors:
            tu, tv, tw = tu.clamp(min=10.0), tv.clamp(min=10.0), tw.clamp(min=30.0)
        ft = self._sample3(self.win["ft"], ft_grid_z)
        above = z > blh
        su, sv, sw = (torch.where(above, ft[0], s) for s in (su, sv, sw))
        tu, tv, tw = (torch.where(above, ft[1], s) for s in (tu, tv, tw))
        return (su.clamp(min=SIGMA_MIN), sv.clamp(min=SIGMA_MIN), sw.clamp(min=SIGMA_MIN),
                tu.clamp(min=T_L_MIN), tv.clamp(min=T_L_MIN), tw.clamp(min=T_L_MIN))

    @staticmethod
    def _ou(prev, tl, s2, dt, eta, drift_dt):
        a = torch.exp(-dt / tl)
        return a * prev + drift_dt + torch.sqrt(torch.clamp(s2 * (1.0 - a * a), min=0.0)) * eta

    def _randn(self, shape):
        return torch.randn(shape, device=self.dev)

    def _rand(self, shape):
        return torch.rand(shape, device=self.dev)

    # -- the step core (hanna.py:898-1068) --------------------------------
    def _turbulence(self, lon0, lat0, z0, up, vp, wp, um, vm, dt, al):
        """One step of the given particles; returns their new state."""

        h = self.h
        W = self.win

        def wind(lo, la, zz):
            v = self._sample3(W["uvw"], self._norm(lo, la, zz))
            v = v[:3] * (1.0 - al) + v[3:] * al
            mlon = M_LON_EQ * torch.clamp(torch.cos(torch.deg2rad(la)).abs(), min=0.05)
            return v[0] / mlon, v[1] / M_LAT, v[2]

        d0, d1, d2 = wind(lon0, lat0, z0)
        d0, d1, d2 = wind(lon0 - 0.5 * dt * d0, lat0 - 0.5 * dt * d1, z0 - 0.5 * dt * d2)
        lon, lat, z = lon0 - dt * d0, lat0 - dt * d1, z0 - dt * d2
        s = self._sample2(torch.lerp(W["surf_s"], W["surf_e"], al), lon, lat)
        blh, sp = torch.clamp(s[0], min=BLH_MIN), s[1]
        ustar, shf, tsf = torch.clamp(s[2], min=USTAR_MIN), s[3], s[4]
        rho_a = sp / (R_DRY * tsf)
        den = KARMAN * GRAVITY * shf
        fin = den.abs() > 1e-10
        Lo = torch.where(fin, -(ustar ** 3) * tsf * rho_a * C_P / torch.where(fin, den, torch.ones_like(den)),
                         torch.full_like(den, math.inf))
        hol = blh / Lo
        cube = GRAVITY * blh * shf / (tsf * rho_a * C_P)
        wstar = torch.where(cube > 0, torch.pow(torch.clamp(cube, min=0.0), 1.0 / 3.0), torch.zeros_like(cube))
        f = 2.0 * OMEGA * torch.sin(torch.deg2rad(lat))
        z_eval = torch.clamp(z, min=h.z_ubl_m)
        g_eval = self._norm(lon, lat, z_eval)
        su, sv, sw, tu, tv, tw = self._column(z_eval, blh, ustar, wstar, hol, f, g_eval)
        z_hi = z_eval + 1.0
        z_lo = torch.clamp(z_eval - 1.0, min=h.z_ubl_m)
        sw_hi = self._column(z_hi, blh, ustar, wstar, hol, f, self._norm(lon, lat, z_hi))[2]
        sw_lo = self._column(z_lo, blh, ustar, wstar, hol, f, self._norm(lon, lat, z_lo))[2]
        dsig2 = (sw_hi ** 2 - sw_lo ** 2) / torch.clamp(z_hi - z_lo, min=1.0)
        sw_sq = torch.clamp(sw ** 2, min=SIGMA_MIN ** 2)
        dens = self._sample3(W["dens"], g_eval)
        dens_b = -(sw_sq * dens[1] / torch.clamp(dens[0], min=1e-3))
        half_b = -0.5 * dsig2
        k = torch.clamp(torch.ceil(dt / (h.substep_c * torch.clamp(tw, min=1e-3))), 1, h.max_substeps)
        sub = dt / k
        # static: the fixed max_substeps loop; dynamic: max(k) iterations (one host read)
        n_loop = int(k.max().item()) if self.dynamic and k.numel() else h.max_substeps
        for i in range(n_loop):
            sdt = torch.where(k > i, sub, torch.zeros_like(sub))
            eta = self._randn((3, lon.numel()))
            drift = (1.0 + wp ** 2 / sw_sq) * half_b + dens_b
            un = self._ou(up, tu, su ** 2, sdt, eta[0], 0.0)
            vn = self._ou(vp, tv, sv ** 2, sdt, eta[1], 0.0)
            wn = self._ou(wp, tw, sw ** 2, sdt, eta[2], drift * sdt)
            wn = torch.clamp(wn, -4.0 * sw, 4.0 * sw)
            un = torch.clamp(un, -4.0 * su, 4.0 * su)
            vn = torch.clamp(vn, -4.0 * sv, 4.0 * sv)
            z = z - wn * sdt
            mlon = M_LON_EQ * torch.clamp(torch.cos(torch.deg2rad(lat)).abs(), min=0.05)
            lon = lon - un * sdt / mlon
            lat = lat - vn * sdt / M_LAT
            below = z < 0.0
            z = torch.where(below, -z, z)
            wn = torch.where(below, -wn, wn)
            up, vp, wp = un, vn, wn
        if h.meander_enabled:
            sm = self._sample3(W["mea"], g_eval)
            eta = self._randn((2, lon.numel()))
            tm = torch.full_like(sub, h.meander_timescale_s)
            um = self._ou(um, tm, sm[0] ** 2, dt, eta[0], 0.0)
            vm = self._ou(vm, tm, sm[1] ** 2, dt, eta[1], 0.0)
            mlon = M_LON_EQ * torch.clamp(torch.cos(torch.deg2rad(lat)).abs(), min=0.05)
            lon = lon - um * dt / mlon
            lat = lat - vm * dt / M_LAT
        return lon, lat, z, up, vp, wp, um, vm

    def _advance(self) -> torch.Tensor:
        """Turbulence step of every active particle (in place); returns the mask."""

        act = (self.r_off >= self.co) & (self.lo_off <= self.co) & self.alive
        st = (self.lon, self.lat, self.z, self.up, self.vp, self.wp, self.um, self.vm)
        if self.dynamic:
            idx = act.nonzero().squeeze(1)
            if idx.numel() == 0:
                return act
            new = self._turbulence(*(a[idx] for a in st), self.dt, self.al)
            for a, b in zip(st, new):
                a.index_copy_(0, idx, b)
        else:
            new = self._turbulence(*st, self.dt, self.al)
            for a, b in zip(st, new):  # the gate (hanna.py:1056-1068), in place
                a.copy_(torch.where(act, b, a))
        return act

    def _grid(self, act) -> None:
        """Gridding (footprint_gridder.py:117-197) and the escape kill (main.py:690-711)."""

        g = self.fg
        x = torch.floor((self.lon - g["lon_min"]) / max(1e-6, g["lon_max"] - g["lon_min"]) * g["n_x"]).long()
        y = torch.floor((self.lat - g["lat_min"]) / max(1e-6, g["lat_max"] - g["lat_min"]) * g["n_y"]).long()
        zb = torch.bucketize(self.z, self.z_edges, right=False) - 1
        tb = torch.floor(torch.clamp(self.we_off - self.co, min=0.0) / 3600.0).long().clamp(max=g["n_t"] - 1)
        valid = act & (x >= 0) & (x < g["n_x"]) & (y >= 0) & (y < g["n_y"]) & (zb >= 0) & (zb < g["n_z"])
        flat = ((((self.rel * g["n_t"] + tb) * g["n_z"] + zb) * g["n_y"] + y) * g["n_x"] + x)
        flat = torch.where(valid, flat, torch.zeros_like(flat))
        self.footprint.scatter_add_(0, flat, self.weight * valid.to(torch.float32) * self.dt)
        dm = self.dom
        inside = ((self.lon >= dm["lon_min"]) & (self.lon <= dm["lon_max"]) & (self.lat >= dm["lat_min"])
                  & (self.lat <= dm["lat_max"]) & (self.z <= dm["alt_max"]))
        self.alive &= ~(act & ~inside)
        self.act_count += act

    def _body(self) -> None:
        self._grid(self._advance())

    # -- convection (emanuel.py:570-805, per column) -----------------------
    def _convect(self, act) -> None:
        """Backward redistribution of the active particles through their
        nearest column's mass-flux matrix (trigger and closure emanuel.py:
        629-679, matrix 429-506, transposed row sampling 720-805)."""

        W = self.win
        lev = self.levels
        Z = lev.numel()
        cape, lcl_h, top_h, dtv = W["conv"]
        ny, nx = cape.shape
        ge_l = lev[:, None, None] >= lcl_h[None]
        ge_t = lev[:, None, None] >= top_h[None]
        none = torch.full((ny, nx), -1, dtype=torch.int64, device=self.dev)
        lcl = torch.where(ge_l.any(0), ge_l.to(torch.int32).argmax(0).long(), none)
        lnb = torch.where(ge_t.any(0), ge_t.to(torch.int32).argmax(0).long(), none.fill_(Z - 1))
        lcl_c = lcl.clamp(min=0)
        fires = (lcl >= 0) & (lnb > lcl) & (cape >= 50.0) & (dtv >= 0.9) & ((lev[lnb] - lev[lcl_c]) >= 500.0)
        cols = fires.reshape(-1).nonzero().squeeze(1)
        if cols.numel() == 0:
            return
        tl = (torch.gather(W["t_s"], 0, lcl_c[None])[0] * (1.0 - self.al)
              + torch.gather(W["t_e"], 0, lcl_c[None])[0] * self.al)
        p = W["p_hpa"].to(torch.float32) * 100.0
        rho = p[lcl_c] / (R_DRY * torch.clamp(tl, min=150.0))
        mb = 0.03 * rho * torch.clamp(torch.sqrt(2.0 * torch.clamp(cape, min=0.0)), max=5.0) * 3600.0
        m = torch.empty_like(p)
        m[1:-1] = 0.5 * (p[:-2] - p[2:]).abs()
        m[0] = 0.5 * (p[0] - p[1]).abs()
        m[-1] = 0.5 * (p[-2] - p[-1]).abs()
        m = m / GRAVITY
        # the firing columns' matrices, materialised [n_fire, Z, Z]
        a, b, mbc = lcl.reshape(-1)[cols], lnb.reshape(-1)[cols], mb.reshape(-1)[cols]
        kk = torch.arange(Z, device=self.dev)
        zero = torch.zeros((), device=self.dev)
        below = kk[None] < a[:, None]
        bl = torch.clamp((m[None] * below).sum(1, keepdim=True), min=1e-12)
        e = torch.where(below, mbc[:, None] * m[None] / bl, zero)
        depth = (b - a).to(torch.float32)[:, None]
        w = torch.where((kk[None] >= a[:, None]) & (kk[None] <= b[:, None]), 1.0 - (kk[None] - a[:, None]) / depth,
                        zero)
        d = mbc[:, None] * w / torch.clamp(w.sum(1, keepdim=True), min=1e-12)
        f = e[:, :, None] * d[:, None, :] / mbc[:, None, None]
        phi = torch.clamp(torch.cumsum(e, 1) - torch.cumsum(d, 1), min=0.0)[:, :-1]
        f = f + torch.diag_embed(phi, offset=-1)
        mx = (f.sum(2) / torch.clamp(m[None], min=1e-12)).amax(1)
        f = f * torch.where(mx > CFL_MAX_OUTFLUX_FRAC, CFL_MAX_OUTFLUX_FRAC / mx, torch.ones_like(mx))[:, None, None]
        lm = torch.clamp(m, min=1e-6)
        ff = f + torch.diag_embed(torch.clamp(lm[None] - f.sum(2), min=0.0))
        prob = torch.clamp(ff.transpose(1, 2) / lm[None, :, None], min=0.0)
        slot = torch.full((ny * nx,), -1, dtype=torch.int64, device=self.dev)
        slot[cols] = torch.arange(cols.numel(), device=self.dev)
        # particles: nearest column, nearest level (emanuel.py:756-787)
        sx = (nx - 1) / (self.lon1 - self.lon0)
        sy = (ny - 1) / (self.lat1 - self.lat0)
        cx = torch.floor(((self.lon - self.lon0) * sx).clamp(0, nx - 1) + 0.5).long().clamp(max=nx - 1)
        cy = torch.floor(((self.lat - self.lat0) * sy).clamp(0, ny - 1) + 0.5).long().clamp(max=ny - 1)
        sl = slot[cy * nx + cx]
        moving = act & (sl >= 0)
        up = torch.searchsorted(lev, self.z.contiguous()).clamp(1, Z - 1)
        host = torch.where((self.z - lev[up - 1]) < (lev[up] - self.z), up - 1, up)
        u = self._rand((2, self.n))
        rows = prob[sl.clamp(min=0), host]
        rows = rows / torch.clamp(rows.sum(1, keepdim=True), min=1e-12)
        dest = (torch.cumsum(rows, 1) < u[0][:, None]).sum(1).clamp(max=Z - 1)
        zl = lev[(dest - 1).clamp(min=0)]
        znew = zl + u[1] * (lev[dest] - zl)
        self.z.copy_(torch.where(moving & (dest != host), znew, self.z))

    @property
    def active_steps(self) -> int:
        """Active particle-steps so far (one reduction, read on the host)."""

        return int(self.act_count.sum().item())

    # -- one step ---------------------------------------------------------
    def step(self) -> None:
        st = self.sched[self.step_i]
        new_hour = self._win_hour != st.bracket_hour
        if new_hour:
            self._load_window(st.bracket_hour)
        self.co.fill_(float(st.cursor_off))
        self.dt.fill_(float(st.delta_s))
        self.al.fill_(float(st.alpha))
        if new_hour and self.convection:
            act = self._advance()  # convection comes after the turbulence step, before the gridding
            self._convect(act)
            self._grid(act)
        elif self.graph and self._warm:  # captured after one eager step has initialised every op
            if self._graph is None:
                self._graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(self._graph):
                    self._body()
            self._graph.replay()
        else:
            self._body()
        self._warm = True
        self.step_i += 1

    def run(self, n_steps: int) -> None:
        for _ in range(min(n_steps, len(self.sched) - self.step_i)):
            self.step()
