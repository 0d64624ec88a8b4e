"""Shared set-ups of the physics acceptance gates (the reference's analytic
dispersion, diffusion-PDE, plume-footprint and well-mixed tests), run through
the engine's fused step (``glide_step`` through the C ABI) and through the
oracle on the same window.

A gate's met is an oracle ``Window`` (numpy), packed for the device in the
documented float32 window layout (``include/glide.h``, GlideWindow), so both
paths read the same fields. Homogeneous or prescribed turbulence uses the
step's GLIDE_F_PRESCRIBED mode: sigma_u = sigma_v = sigma_w and T_L from the
support record at every height, which is how the reference's engine-level
tests drive the OU primitives with given (sigma_w, T_L)."""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from oracle import lpdm as O

M_PER_DEG_LAT = 110540.0
M_PER_DEG_LON = 111320.0


def taylor_var(sigma_w2: float, t_l: float, t: float) -> float:
    """Taylor (1921) displacement variance of an OU velocity process."""
    return 2.0 * sigma_w2 * t_l * (t - t_l * (1.0 - math.exp(-t / t_l)))


def uniform_window(*, levels, lon, lat, sigma=None, t_l=None, u=0.0, v=0.0, wind_fn=None, blh=100.0,
                   ustar=0.3, shf=0.0, t_k=280.0, sp=101325.0) -> O.Window:
    """A window on the grid (lon [X], lat [Y], levels [Z]): winds u, v (or
    wind_fn(LON, LAT) -> (u, v) on the grid), w = 0, both bracket ends equal;
    support stack sigma (scalar or [Z] profile), T_L (scalar or [Z]); flat
    density (rho 1, d rho / dz 0); no meander; uniform surface fields."""

    levels = np.asarray(levels, np.float64)
    lon = np.asarray(lon, np.float64)
    lat = np.asarray(lat, np.float64)
    Z, Y, X = levels.size, lat.size, lon.size
    shape = (Z, Y, X)
    uvw = np.zeros((3,) + shape)
    if wind_fn is not None:
        LAT, LON = np.meshgrid(lat, lon, indexing="ij")
        uu, vv = wind_fn(LON, LAT)
        uvw[0] = uu[None]
        uvw[1] = vv[None]
    else:
        uvw[0] = u
        uvw[1] = v
    ft = np.zeros((2,) + shape)
    ft[0] = np.broadcast_to(np.asarray(sigma if sigma is not None else 0.1, np.float64).reshape(-1, 1, 1), shape)
    ft[1] = np.broadcast_to(np.asarray(t_l if t_l is not None else 100.0, np.float64).reshape(-1, 1, 1), shape)
    dens = np.zeros((2,) + shape)
    dens[0] = 1.0
    surf = np.zeros((5, Y, X))
    surf[0], surf[1], surf[2], surf[3], surf[4] = blh, sp, ustar, shf, t_k
    return O.Window(uvw_s=uvw, uvw_e=uvw.copy(), surf_s=surf, surf_e=surf.copy(), ft=ft, dens=dens,
                    meander=np.zeros((2,) + shape), lon0=float(lon[0]), lon1=float(lon[-1]), lat0=float(lat[0]),
                    lat1=float(lat[-1]), levels=levels)


def pack_window(win: O.Window):
    """float32 packed window (wind, aux, surf as uint8 numpy arrays): cells
    column-major (y, x, level); wind plane A = u_s v_s u_e v_e, plane B = w_s
    w_e; support plane A = sigma T_L rho drho_dz, plane B = sigma_mu sigma_mv;
    surface records blh sp ustar shf t at both ends, then two zeros."""

    def cm(a):  # [Z, Y, X] -> cells (y, x, z)
        return np.ascontiguousarray(np.moveaxis(np.asarray(a, np.float32), 0, -1)).reshape(-1)

    us, vs, ws = win.uvw_s
    ue, ve, we = win.uvw_e
    wind = np.concatenate([np.stack([cm(us), cm(vs), cm(ue), cm(ve)], 1).reshape(-1),
                           np.stack([cm(ws), cm(we)], 1).reshape(-1)]).astype(np.float32)
    aux = np.concatenate([np.stack([cm(win.ft[0]), cm(win.ft[1]), cm(win.dens[0]), cm(win.dens[1])], 1).reshape(-1),
                          np.stack([cm(win.meander[0]), cm(win.meander[1])], 1).reshape(-1)]).astype(np.float32)
    Y, X = win.surf_s.shape[1:]
    surf = np.zeros((Y, X, 12), np.float32)
    surf[..., 0:5] = np.moveaxis(win.surf_s, 0, -1)
    surf[..., 5:10] = np.moveaxis(win.surf_e, 0, -1)
    return wind.view(np.uint8), aux.view(np.uint8), surf.reshape(-1).view(np.uint8)


class GpuStepper:
    """Particles in the reference's (N, 4) [lon, lat, alt, weight] layout
    (float32, on the GPU) advanced by ``glide_step`` over one packed window,
    every particle active (explicit mask). ``state`` holds u', v', w' (and the
    meander pair)."""

    def __init__(self, win: O.Window, lon, lat, z, *, flags: int, seed: int, w0=None, u0=None, v0=None,
                 substep_c: float = 0.5, max_substeps: int = 6, z_ubl_m: float = 2.0):
        import torch

        from glide_amd import _lib as L

        self.L, self.torch = L, torch
        n = int(np.asarray(z).size)
        self.n = n
        P = np.zeros((n, 4), np.float32)
        P[:, 0] = lon
        P[:, 1] = lat
        P[:, 2] = z
        P[:, 3] = 1.0 / n
        self.P = torch.from_numpy(P).cuda()
        zeros = lambda: torch.zeros(n, dtype=torch.float32, device="cuda")  # noqa: E731
        self.state = {k: zeros() for k in ("u_prime", "v_prime", "w_prime", "u_meander", "v_meander")}
        for k, v0_ in (("w_prime", w0), ("u_prime", u0), ("v_prime", v0)):
            if v0_ is not None:
                self.state[k].copy_(torch.from_numpy(np.asarray(v0_, np.float32)))
        self.uid = torch.arange(n, dtype=torch.int32, device="cuda")
        self.active = torch.ones(n, dtype=torch.uint8, device="cuda")
        wind, aux, surf = pack_window(win)
        self.buf = [torch.from_numpy(a.copy()).cuda() for a in (wind, aux, surf)]
        self.levels = torch.from_numpy(np.asarray(win.levels, np.float32)).cuda()
        Y, X = win.surf_s.shape[1:]
        self.grid = L.GlideGrid(nx=X, ny=Y, nz=int(self.levels.numel()), reserved=0, lon0=win.lon0, lon1=win.lon1,
                                lat0=win.lat0, lat1=win.lat1, levels=self.levels.data_ptr())
        self.win = L.GlideWindow(wind=self.buf[0].data_ptr(), aux=self.buf[1].data_ptr(), surf=self.buf[2].data_ptr(),
                                 dtype=L.GLIDE_MET_F32, reserved=0)
        self.flags = flags | L.F_MASK_INPUT
        self.seed = int(seed)
        self.substep_c, self.max_substeps, self.z_ubl_m = substep_c, max_substeps, z_ubl_m
        self.k = 0

    def parts(self):
        L, s, base = self.L, self.state, self.P.data_ptr()
        return L.GlideParticles(
            n=self.n, pos_stride=4, lon=base, lat=base + 4, z=base + 8, up=s["u_prime"].data_ptr(),
            vp=s["v_prime"].data_ptr(), wp=s["w_prime"].data_ptr(), um=s["u_meander"].data_ptr(),
            vm=s["v_meander"].data_ptr(), weight=base + 12, r_off=None, we_off=None, rel_idx=None,
            uid=self.uid.data_ptr(), alive=None, active=self.active.data_ptr())

    def step(self, dt: float, alpha: float = 0.5) -> None:
        L = self.L
        p = L.GlideStepParams(dt=float(dt), alpha=float(alpha), cursor_off=0.0, length_s=0.0,
                              substep_c=self.substep_c, z_ubl_m=self.z_ubl_m, meander_timescale_s=1800.0,
                              seed=self.seed, step_index=self.k, max_substeps=self.max_substeps,
                              flags=self.flags, block=0)
        parts = self.parts()
        L.check(L.lib().glide_step(C.byref(p), C.byref(parts), C.byref(self.grid), C.byref(self.win), None, None,
                                   L.stream_ptr()))
        self.k += 1

    @property
    def z(self):
        return self.P[:, 2]


class OracleStepper:
    """The same particles advanced by the float64 oracle step on the same window."""

    def __init__(self, win: O.Window, lon, lat, z, *, params: O.HannaParams, seed: int, w0=None, u0=None, v0=None):
        n = int(np.asarray(z).size)
        self.win = win
        self.params = params
        self.seed = int(seed)
        f = np.float64
        self.S = dict(lon=np.broadcast_to(np.asarray(lon, f), (n,)).copy(),
                      lat=np.broadcast_to(np.asarray(lat, f), (n,)).copy(), z=np.asarray(z, f).copy(),
                      up=np.zeros(n) if u0 is None else np.asarray(u0, f).copy(),
                      vp=np.zeros(n) if v0 is None else np.asarray(v0, f).copy(),
                      wp=np.zeros(n) if w0 is None else np.asarray(w0, f).copy(),
                      um=np.zeros(n), vm=np.zeros(n), uid=np.arange(n, dtype=np.uint32))
        self.active = np.ones(n, bool)
        self.k = 0

    def step(self, dt: float, alpha: float = 0.5) -> None:
        self.S = O.step(self.S, self.active, self.win, alpha, float(dt), self.params, seed=self.seed,
                        step_index=self.k)
        self.k += 1


# ---------------------------------------------------------------------------
# Solid-body rotation near the pole, where the step's metric
# 111320 * max(|cos(lat)|, 0.05) is the constant 5566 m per degree of
# longitude: u = -omega Y, v = omega X (X, Y metres from the centre) is then a
# rigid rotation of the particles' (X, Y), and linear in (lon, lat), so the
# trilinear sampling reproduces it exactly.
# ---------------------------------------------------------------------------
POLE_LAT_C = 88.95
POLE_M_PER_DEG_LON = M_PER_DEG_LON * 0.05


def rotation_window(omega: float) -> O.Window:
    lon = np.arange(-20.0, 20.0 + 1e-9, 0.5)
    lat = POLE_LAT_C + 1.0 - 0.05 * np.arange(41)  # 89.95 .. 87.95 (descending)

    def wind(LON, LAT):
        return -omega * (LAT - POLE_LAT_C) * M_PER_DEG_LAT, omega * LON * POLE_M_PER_DEG_LON

    return uniform_window(levels=[0.0, 20000.0], lon=lon, lat=lat, wind_fn=wind, blh=100.0)


def rotation_start():
    xy = np.array([[100_000.0, 0.0], [0.0, 50_000.0], [-70_000.0, 40_000.0], [30_000.0, -80_000.0]])
    return xy[:, 0] / POLE_M_PER_DEG_LON, POLE_LAT_C + xy[:, 1] / M_PER_DEG_LAT, xy


def rotation_xy(lon, lat):
    return np.stack([np.asarray(lon, np.float64) * POLE_M_PER_DEG_LON,
                     (np.asarray(lat, np.float64) - POLE_LAT_C) * M_PER_DEG_LAT], 1)


# ---------------------------------------------------------------------------
# Diffusion PDE (reference tests/test_diffusion_pde.py): K(z) = K0 + a min(z,
# z_cap), constant T_L, release slab inside the low-K layer, conservative
# flux-form Crank-Nicolson with zero-flux boundaries.
# ---------------------------------------------------------------------------
PDE_K0, PDE_A, PDE_ZCAP, PDE_HTOP, PDE_TL, PDE_DT = 0.02, 0.12, 200.0, 400.0, 10.0, 1.0
PDE_Z0 = (5.0, 15.0)
PDE_BIN_W, PDE_DZ = 20.0, 2.0


def pde_k_faces(z_faces: np.ndarray, near_surface_factor: float = 1.0) -> np.ndarray:
    k = PDE_K0 + PDE_A * np.minimum(z_faces, PDE_ZCAP)
    k = k.copy()
    k[z_faces < 50.0] *= near_surface_factor
    return k


def pde_stepper(k_faces: np.ndarray, n_nodes: int):
    from scipy.linalg import solve_banded

    main = np.zeros(n_nodes)
    main[0] = k_faces[0]
    main[1:-1] = k_faces[:-1] + k_faces[1:]
    main[-1] = k_faces[-1]
    lam = PDE_DT / PDE_DZ**2
    ab = np.zeros((3, n_nodes))
    ab[0, 1:] = -0.5 * lam * k_faces
    ab[1, :] = 1.0 + 0.5 * lam * main
    ab[2, :-1] = -0.5 * lam * k_faces

    def step(c):
        rhs = c - 0.5 * lam * (main * c - np.r_[0.0, k_faces * c[:-1]] - np.r_[k_faces * c[1:], 0.0])
        return solve_banded((1, 1), ab, rhs)

    return step


def pde_levels() -> np.ndarray:
    """A fine ladder for the prescribed sigma(z) profile: 0.25 m first layer,
    stretched to 600 m (above the 400 m reflecting top), 128 levels, with the
    cap height on a level."""
    from glide_amd.synthetic import stretched_agl_levels

    lev = stretched_agl_levels(127, 600.0, 0.25)
    return np.sort(np.unique(np.concatenate([lev, [PDE_ZCAP]])))


def pde_sigma(levels: np.ndarray) -> np.ndarray:
    return np.sqrt((PDE_K0 + PDE_A * np.minimum(levels, PDE_ZCAP)) / PDE_TL)


def pde_reference(checkpoints):
    """Crank-Nicolson solutions (true / half / quarter near-surface K) binned
    at the checkpoints."""

    z_nodes = np.arange(0.0, PDE_HTOP + PDE_DZ / 2, PDE_DZ)
    z_faces = (z_nodes[:-1] + z_nodes[1:]) / 2
    steppers = {name: pde_stepper(pde_k_faces(z_faces, f), len(z_nodes))
                for name, f in (("true", 1.0), ("half", 0.5), ("quarter", 0.25))}
    c0 = np.where((z_nodes >= PDE_Z0[0]) & (z_nodes <= PDE_Z0[1]), 1.0, 0.0)
    c0 /= c0.sum()
    pde = {name: c0.copy() for name in steppers}
    edges = np.arange(0.0, PDE_HTOP + PDE_BIN_W / 2, PDE_BIN_W)
    node_bin = np.minimum((z_nodes // PDE_BIN_W).astype(int), len(edges) - 2)
    out = {}
    t = 0.0
    for tc in checkpoints:
        for _ in range(int(round((tc - t) / PDE_DT))):
            for name in pde:
                pde[name] = steppers[name](pde[name])
        t = tc
        out[tc] = {}
        for name, c in pde.items():
            h = np.bincount(node_bin, weights=c, minlength=len(edges) - 1)
            out[tc][name] = h / h.sum()
    return out, edges


# ---------------------------------------------------------------------------
# Gaussian plume (reference tests/test_plume_footprint.py)
# ---------------------------------------------------------------------------
PL_U, PL_SIGMA2, PL_TL, PL_ZR, PL_HSURF, PL_DT = 5.0, 0.25, 100.0, 50.0, 40.0, 5.0
PL_DX, PL_DY, PL_NX, PL_NY = 250.0, 50.0, 82, 100
PL_X0, PL_Y0 = -20250.0, -2500.0


def plume_taylor_sigma(t: float) -> float:
    return math.sqrt(2.0 * PL_SIGMA2 * PL_TL * (t - PL_TL * (1.0 - math.exp(-t / PL_TL))))


def plume_window() -> O.Window:
    lon = np.linspace(-0.2, 0.02, 4)
    lat = np.linspace(0.03, -0.03, 4)
    return uniform_window(levels=[0.0, 20000.0], lon=lon, lat=lat, sigma=math.sqrt(PL_SIGMA2), t_l=PL_TL,
                          u=PL_U, blh=100.0)


def plume_analytic(n_steps: int) -> dict:
    from scipy.special import ndtr

    x_c = (np.arange(PL_NX) + 0.5) * PL_DX + PL_X0
    y_c = (np.arange(PL_NY) + 0.5) * PL_DY + PL_Y0
    y_e = np.arange(PL_NY + 1) * PL_DY + PL_Y0
    analytic = np.zeros((PL_NY, PL_NX))
    for j in range(PL_NX):
        if x_c[j] >= 0.0:
            continue
        t = -x_c[j] / PL_U
        s = plume_taylor_sigma(t)
        p_y = ndtr(y_e[1:] / s) - ndtr(y_e[:-1] / s)
        p_z = (ndtr((PL_HSURF - PL_ZR) / s) - ndtr(-PL_ZR / s)) + (ndtr((PL_HSURF + PL_ZR) / s) - ndtr(PL_ZR / s))
        analytic[:, j] = (PL_DX / PL_U) * p_y * p_z
    return dict(analytic=analytic, x_centres=x_c, y_centres=y_c, travel_t=-x_c / PL_U, t_max=n_steps * PL_DT)


def plume_checks(raw: np.ndarray, ref: dict, *, bands, corr_min: float, ratio_tol: float, width_tol: float,
                 tl: float = PL_TL) -> None:
    """The reference's plume assertions on a raw surface footprint [NY, NX]."""

    tt = ref["travel_t"]

    def band(lo, hi):
        return np.where((tt >= lo) & (tt <= hi))[0]

    for lo, hi, tol_max, tol_mean in bands:
        cols = band(lo, hi)
        got = raw[:, cols].sum(axis=0)
        want = ref["analytic"][:, cols].sum(axis=0)
        rel = np.abs(got - want) / want
        assert rel.max() < tol_max, (lo, hi, float(rel.max()))
        assert rel.mean() < tol_mean, (lo, hi, float(rel.mean()))
    cols = band(2 * tl, ref["t_max"])
    g, a = raw[:, cols].ravel(), ref["analytic"][:, cols].ravel()
    corr = float(np.corrcoef(g, a)[0, 1])
    assert corr > corr_min, corr
    ratio = raw[:, cols].sum() / ref["analytic"][:, cols].sum()
    assert 1.0 - ratio_tol < ratio < 1.0 + ratio_tol, ratio
    yc = ref["y_centres"]
    for c in cols[:: max(1, len(cols) // 6)]:
        col = raw[:, c]
        assert col.sum() > 0
        mu = float((col * yc).sum() / col.sum())
        var = float((col * (yc - mu) ** 2).sum() / col.sum()) - PL_DY**2 / 12.0
        got = math.sqrt(max(var, 1e-9))
        want = plume_taylor_sigma(tt[c])
        assert abs(got - want) / want < width_tol, (float(tt[c]), got, want)


def plume_gridder_bounds():
    return dict(lon_bounds=(PL_X0 / M_PER_DEG_LON, (PL_X0 + PL_NX * PL_DX) / M_PER_DEG_LON),
                lat_bounds=(PL_Y0 / M_PER_DEG_LAT, (PL_Y0 + PL_NY * PL_DY) / M_PER_DEG_LAT),
                z_edges_m=(0.0, PL_HSURF, 10000.0))


# ---------------------------------------------------------------------------
# Well-mixed (reference tests/test_main_runtime.py:573-921)
# ---------------------------------------------------------------------------
def wmc_met(*, blh_m: float, ustar_m_s: float = 0.4, p_top_hpa: float = 600.0, z_top: float | None = None,
            n_lev: int = 6, n_lat: int = 4, n_lon: int = 4):
    """Spatially uniform neutral BL met window (reference _wmc_met_window and
    _wmc_density_met_window) as HourlyMetTensors of CPU tensors."""

    import datetime as _dt

    import torch

    from glide_amd.met import HourlyMetTensors, MetFieldMetadata

    shape = (n_lev, n_lat, n_lon)
    zeros = torch.zeros(shape, dtype=torch.float32)
    chans = {"u": zeros, "v": zeros, "w": zeros, "blh": torch.full(shape, blh_m), "sp": torch.full(shape, 101325.0),
             "t": torch.full(shape, 280.0), "ustar": torch.full(shape, ustar_m_s), "shf": torch.zeros(shape)}
    names = ("u", "v", "w", "blh", "sp", "t", "ustar", "shf")
    fields = torch.stack([chans[n] for n in names], dim=0)
    t0 = _dt.datetime(2024, 1, 1, 12, 0, 0, tzinfo=_dt.timezone.utc)
    level = np.linspace(0.0, z_top if z_top is not None else max(4000.0, 1.5 * blh_m), n_lev)
    md = MetFieldMetadata(lon=np.linspace(-2.0, 2.0, n_lon), lat=np.linspace(40.0, 50.0, n_lat), level=level,
                          pressure_level_hpa=np.linspace(1000.0, p_top_hpa, n_lev), time_start=t0,
                          time_end=t0 + _dt.timedelta(hours=1), variable_units=dict.fromkeys(names, ""))
    height = torch.as_tensor(level, dtype=torch.float32).view(n_lev, 1, 1).expand(shape).contiguous()
    return HourlyMetTensors(hour_start=fields, hour_end=fields, metadata=md, channel_names=names,
                            height_agl_m=height)


def wmc_oracle_window(met) -> O.Window:
    """The oracle's window of a well-mixed met (the same construction as a
    resampled hour: winds, t, q = 0 on the levels; surface fields)."""

    from oracle import run as ORUN

    hs = met.hour_start.numpy().astype(np.float64)
    names = met.channel_names
    ix = {n: i for i, n in enumerate(names)}
    Z, Y, X = hs.shape[1:]
    out3 = np.stack([hs[ix["u"]], hs[ix["v"]], hs[ix["w"]], hs[ix["t"]], np.zeros((Z, Y, X))])
    src2 = np.stack([hs[ix[c]][0] for c in ("blh", "sp", "ustar", "shf")])
    md = met.metadata
    hour = (out3, src2, np.asarray(md.pressure_level_hpa, np.float64))
    return ORUN.build_window(hour, hour, md.level, md.lon, md.lat, np.float64)


def density_at(z_m, *, p_bot_hpa, p_top_hpa, z_top_m, t_kelvin):
    p_hpa = p_bot_hpa + (p_top_hpa - p_bot_hpa) * (np.asarray(z_m) / z_top_m)
    return (p_hpa * 100.0) / (287.05 * t_kelvin)
