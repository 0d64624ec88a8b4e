"""The plain FP32 method as PyTorch ops on the GPU: the comparison baseline.

This is the reference's own GPU path for one step, restated with the same
torch operations it uses (HannaScheme static path, hanna.py:855-1068;
GPUEngine.normalize_particle_coordinates, gpu_engine.py:210-299; the driver's
mask, gridding and escape kill, main.py:1396-1477, footprint_gridder.py:
117-197): trilinear and bilinear ``grid_sample(align_corners=True)``, a
``searchsorted`` level lookup, ``torch.randn`` noise, the fixed
``max_substeps`` loop with finished particles advanced by ``sub_dt = 0``, and
``scatter_add_`` gridding with float atomics. Every operation is a PyTorch
kernel; none of this build's HIP kernels run inside a step. The hourly window
(winds of both bracket hours, surface fields, the support stacks) is the one
``glide_build_window`` packs, viewed as separate fields: it is built once per
met hour, as the reference builds its static buffers once per window.

It exists to measure the speedup of the HIP engine over the plain method on
the same GPU; it is not a product path and not a correctness oracle (its noise
is torch's, not the engine's Philox stream). Convection is not applied (one
event per met hour; leaving it out only makes this baseline faster).
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


class TorchPlain:
    """One release batch stepped with PyTorch ops, reusing an engine's met
    (``DeviceMet``), schedule, release arrays and footprint grid."""

    def __init__(self, eng, *, seed: int = 0):
        if eng.met.met_dtype != L.GLIDE_MET_F32:
            raise ValueError("the plain method reads float32 met")
        self.eng = eng
        self.dev = eng.device
        self.h = eng.h
        self.gen = torch.Generator(device=self.dev)
        self.gen.manual_seed(int(seed))
        ra = eng.ra
        d = self.dev
        t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(d, dt)  # noqa: E731
        self.lon, self.lat, self.z = t(ra.lon), t(ra.lat), t(ra.z)
        self.weight, self.r_off, self.we_off = t(ra.weight), t(ra.r_off), t(ra.we_off)
        self.lo_off = t(ra.lo_off)
        self.rel = t(ra.rel_idx, torch.int64)
        n = self.lon.numel()
        self.up, self.vp, self.wp, self.um, self.vm = (torch.zeros(n, device=d) for _ in range(5))
        self.alive = torch.ones(n, dtype=torch.bool, device=d)
        self.footprint = torch.zeros(eng.n_cells, device=d)
        self.levels = eng.met.levels_f
        self.step_i = 0
        self.active_steps = torch.zeros((), dtype=torch.int64, device=d)
        g = eng.grid_cfg
        self.fg = g
        self.z_edges = eng.z_edges
        self.dom = eng.cfg["domain"]
        self._win_hour = None

    # -- met --------------------------------------------------------------
    def _window(self, stream: int) -> None:
        m = self.eng.met
        nz, ny, nx = m.nz, m.ny, m.nx
        # packed column-major window -> [nz][ny][nx] fields
        wind, aux = L.unpack_window(m.wind, m.aux, nz, ny, nx, m.met_dtype)
        surf = m.surf.view(torch.float32).view(ny, nx, 12)
        self.uvw = wind[..., :6].permute(3, 0, 1, 2).unsqueeze(0).contiguous()  # [1, 6, Z, Y, X]
        self.ft = aux[..., 0:2].permute(3, 0, 1, 2).unsqueeze(0).contiguous()
        self.dens = aux[..., 2:4].permute(3, 0, 1, 2).unsqueeze(0).contiguous()
        self.mea = aux[..., 4:6].permute(3, 0, 1, 2).unsqueeze(0).contiguous()
        self.surf_s = surf[..., 0:5].permute(2, 0, 1).contiguous()
        self.surf_e = surf[..., 5:10].permute(2, 0, 1).contiguous()
        g = m.grid
        self.lon0, self.lon1, self.lat0, self.lat1 = g.lon0, g.lon1, g.lat0, g.lat1

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
        out = F.grid_sample(field, grid.view(1, -1, 1, 1, 3), mode="bilinear", padding_mode="zeros",
                            align_corners=True)
        return out.view(field.shape[1], -1)

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
        ft = self._sample3(self.ft, ft_grid_z)
        above = z > blh
        su, sv, sw = (torch.where(above, ft[0], s) for s in (su, sv, sw))
        tu, tv, tw = (torch.where(above, ft[1], s) for s in (tu, tv, tw))
        return (su.clamp(min=SIGMA_MIN), sv.clamp(min=SIGMA_MIN), sw.clamp(min=SIGMA_MIN),
                tu.clamp(min=T_L_MIN), tv.clamp(min=T_L_MIN), tw.clamp(min=T_L_MIN))

    @staticmethod
    def _ou(prev, tl, s2, dt, eta, drift_dt):
        r = -dt / tl
        a = torch.exp(r) if torch.is_tensor(r) else math.exp(r)
        return a * prev + drift_dt + torch.sqrt(torch.clamp(s2 * (1.0 - a * a), min=0.0)) * eta

    # -- one step ---------------------------------------------------------
    def step(self, stream: int) -> None:
        eng = self.eng
        st = eng.sched[self.step_i]
        eng.met.build_window(st.bracket_hour, stream)
        if self._win_hour != st.bracket_hour:
            self._window(stream)
            self._win_hour = st.bracket_hour
        co = float(st.cursor_off)
        dt, al = float(st.delta_s), float(st.alpha)
        act = (self.r_off >= co) & (self.lo_off <= co) & self.alive
        lon0, lat0, z0 = self.lon, self.lat, self.z
        # RK2 midpoint advection (hanna.py:855-896)
        def wind(lo, la, zz):
            v = self._sample3(self.uvw, self._norm(lo, la, zz))
            v = v[:3] * (1.0 - al) + v[3:] * al
            mlon = M_LON_EQ * torch.clamp(torch.cos(torch.deg2rad(la)).abs(), min=0.05)
            return v[0] / mlon, v[1] / M_LAT, v[2]
        d0, d1, d2 = wind(lon0, lat0, z0)
        d0, d1, d2 = wind(lon0 - 0.5 * dt * d0, lat0 - 0.5 * dt * d1, z0 - 0.5 * dt * d2)
        lon, lat, z = lon0 - dt * d0, lat0 - dt * d1, z0 - dt * d2
        # surface scalars lerped on the grid, sampled at the new position
        s = self._sample2(torch.lerp(self.surf_s, self.surf_e, al), lon, lat)
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
        z_eval = torch.clamp(z, min=self.h.z_ubl_m)
        g_eval = self._norm(lon, lat, z_eval)
        su, sv, sw, tu, tv, tw = self._column(z_eval, blh, ustar, wstar, hol, f, g_eval)
        z_hi = z_eval + 1.0
        z_lo = torch.clamp(z_eval - 1.0, min=self.h.z_ubl_m)
        sw_hi = self._column(z_hi, blh, ustar, wstar, hol, f, self._norm(lon, lat, z_hi))[2]
        sw_lo = self._column(z_lo, blh, ustar, wstar, hol, f, self._norm(lon, lat, z_lo))[2]
        dsig2 = (sw_hi ** 2 - sw_lo ** 2) / torch.clamp(z_hi - z_lo, min=1.0)
        sw_sq = torch.clamp(sw ** 2, min=SIGMA_MIN ** 2)
        dens = self._sample3(self.dens, g_eval)
        dens_b = -(sw_sq * dens[1] / torch.clamp(dens[0], min=1e-3))
        half_b = -0.5 * dsig2
        k = torch.clamp(torch.ceil(dt / (self.h.substep_c * torch.clamp(tw, min=1e-3))), 1, self.h.max_substeps)
        sub = torch.where(act, dt / k, torch.zeros_like(k))
        up, vp, wp = self.up, self.vp, self.wp
        for i in range(self.h.max_substeps):
            sdt = torch.where(k > i, sub, torch.zeros_like(sub))
            eta = torch.randn((3, lon.numel()), device=self.dev, generator=self.gen)
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
        um, vm = self.um, self.vm
        if self.h.meander_enabled:
            sm = self._sample3(self.mea, g_eval)
            eta = torch.randn((2, lon.numel()), device=self.dev, generator=self.gen)
            tm = self.h.meander_timescale_s
            um = self._ou(um, tm, sm[0] ** 2, dt, eta[0], 0.0)
            vm = self._ou(vm, tm, sm[1] ** 2, dt, eta[1], 0.0)
            mlon = M_LON_EQ * torch.clamp(torch.cos(torch.deg2rad(lat)).abs(), min=0.05)
            lon = lon - um * dt / mlon
            lat = lat - vm * dt / M_LAT
        # gate (hanna.py:1056-1068)
        self.lon, self.lat, self.z = (torch.where(act, a, b) for a, b in ((lon, lon0), (lat, lat0), (z, z0)))
        self.up, self.vp, self.wp = (torch.where(act, a, b) for a, b in ((up, self.up), (vp, self.vp), (wp, self.wp)))
        self.um, self.vm = (torch.where(act, a, b) for a, b in ((um, self.um), (vm, self.vm)))
        # gridding (footprint_gridder.py:117-197) and the escape kill (main.py:690-711)
        g = self.fg
        x = torch.floor((self.lon - g["lon_min"]) / max(1e-6, g["lon_max"] - g["lon_min"]) * g["n_x"]).long()
        y = torch.floor((self.lat - g["lat_min"]) / max(1e-6, g["lat_max"] - g["lat_min"]) * g["n_y"]).long()
        zb = torch.bucketize(self.z, self.z_edges, right=False) - 1
        tb = torch.floor(torch.clamp(self.we_off - co, min=0.0) / 3600.0).long().clamp(max=g["n_t"] - 1)
        valid = act & (x >= 0) & (x < g["n_x"]) & (y >= 0) & (y < g["n_y"]) & (zb >= 0) & (zb < g["n_z"])
        flat = ((((self.rel * g["n_t"] + tb) * g["n_z"] + zb) * g["n_y"] + y) * g["n_x"] + x)
        flat = torch.where(valid, flat, torch.zeros_like(flat))
        self.footprint.scatter_add_(0, flat, self.weight * valid.to(torch.float32) * dt)
        dm = self.dom
        inside = ((self.lon >= dm["lon_min"]) & (self.lon <= dm["lon_max"]) & (self.lat >= dm["lat_min"])
                  & (self.lat <= dm["lat_max"]) & (self.z <= dm["alt_max"]))
        self.alive = self.alive & ~(act & ~inside)
        self.active_steps += act.sum()  # device counter, read once at the end
        self.step_i += 1

    def run(self, n_steps: int, stream: int) -> None:
        for _ in range(min(n_steps, len(self.eng.sched) - self.step_i)):
            self.step(stream)
