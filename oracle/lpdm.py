"""Numpy restatement of the reference's per-step backward-LPDM body (static path).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

dtype-parametrised: every function computes in the dtype of its inputs (float64
for the oracle, float32 for the CPU baseline). Noise comes from the counter-based
Philox stream in ``oracle/philox.py``, so the static path (fixed ``max_substeps``
iterations, finished particles advanced with ``sub_dt = 0``) and the per-particle
``k_i`` loop coincide exactly: a ``sub_dt = 0`` substep is an identity
(reference ``hanna.py:1245-1254``) and each particle's draws do not depend on
which other particles still owe a substep.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import philox

# Physical constants (reference src/lpdm/turbulence/hanna.py:47-134).
KARMAN = 0.4
GRAVITY = 9.80665
R_DRY = 287.05
C_P = 1005.0
OMEGA_EARTH = 7.2921e-5
SIGMA_MIN = 1e-3
T_L_MIN = 1.0
USTAR_MIN = 1e-3
BLH_MIN = 50.0
Z_MIN = 1.0
T_LU_FLOOR = 10.0
T_LV_FLOOR = 10.0
T_LW_FLOOR = 30.0
SURFACE_LAYER_FRACTION = 0.1
W_PRIME_SIGMA_RATIO_MAX = 4.0
DRIFT_FD_DELTA_M = 1.0
M_PER_DEG_LAT = 110540.0
M_PER_DEG_LON_EQ = 111320.0


@dataclass
class HannaParams:
    """HannaScheme constructor options (reference hanna.py:440-520)."""

    meander_enabled: bool = True
    meander_timescale_s: float = 1800.0
    substep_c: float = 0.5
    max_substeps: int = 6
    z_ubl_m: float = 2.0
    flexpart_tl_floors: bool = True
    surface_layer_override: bool = False
    # prescribed turbulence: the free-troposphere branch (sigma and T_L from the
    # support stack) at every height (the engine's GLIDE_F_PRESCRIBED)
    prescribed: bool = False


@dataclass
class Window:
    """One met window, as the static path's ``_gather_static_inputs`` sees it
    (reference hanna.py:798-853): raw winds at the bracket ends, the five surface
    fields at the bracket ends, and the window-midpoint support stacks."""

    uvw_s: np.ndarray  # [3, Z, Y, X]
    uvw_e: np.ndarray  # [3, Z, Y, X]
    surf_s: np.ndarray  # [5, Y, X]  blh, sp, ustar, shf, t(lowest level)
    surf_e: np.ndarray  # [5, Y, X]
    ft: np.ndarray  # [2, Z, Y, X]  sigma_w_ft (= sigma_uv_ft), T_Lw_ft (= T_Luv_ft)
    dens: np.ndarray  # [2, Z, Y, X]  rho, drho/dz
    meander: np.ndarray  # [2, Z, Y, X]  sigma_mu, sigma_mv
    lon0: float
    lon1: float
    lat0: float
    lat1: float
    levels: np.ndarray  # [Z] per-level AGL heights (ascending)

    def astype(self, dt) -> "Window":
        c = lambda a: np.ascontiguousarray(a, dtype=dt)  # noqa: E731
        return Window(
            c(self.uvw_s), c(self.uvw_e), c(self.surf_s), c(self.surf_e), c(self.ft),
            c(self.dens), c(self.meander), float(self.lon0), float(self.lon1),
            float(self.lat0), float(self.lat1), c(self.levels),
        )


# ---------------------------------------------------------------------------
# Interpolation (torch grid_sample, align_corners=True, restated)
# ---------------------------------------------------------------------------


def torch_lerp(a, b, w):
    """torch.lerp(start, end, weight) for scalar ``w`` (ATen lerp: the two-sided form)."""

    if abs(w) < 0.5:
        return a + w * (b - a)
    return b - (b - a) * (1.0 - w)


def level_fraction(z, levels):
    """Fractional (ascending) level index of ``z``: reference
    gpu_engine.py:264-292 (searchsorted right=False, idx clamped to [1, n-1],
    frac = clamp((z - L_lo)/max(L_hi - L_lo, 1e-6), 0, 1))."""

    n = levels.shape[0]
    idx_up = np.clip(np.searchsorted(levels, z, side="left"), 1, n - 1)
    idx_lo = idx_up - 1
    lo = levels[idx_lo]
    hi = levels[idx_up]
    frac = (z - lo) / np.maximum(hi - lo, z.dtype.type(1e-6))
    return idx_lo.astype(z.dtype) + np.clip(frac, 0.0, 1.0)


def normalize(lon, lat, z, w: Window):
    """Reference gpu_engine.py:210-299 with ``level_arr`` ascending: returns the
    grid_sample coordinates (x, y, z) in [-1, 1]."""

    x = 2.0 * ((lon - w.lon0) / (w.lon1 - w.lon0)) - 1.0
    y = 2.0 * ((lat - w.lat0) / (w.lat1 - w.lat0)) - 1.0
    n = w.levels.shape[0]
    fi = level_fraction(z, w.levels)
    zn = 2.0 * fi / float(n - 1) - 1.0
    return np.clip(x, -1.0, 1.0), np.clip(y, -1.0, 1.0), np.clip(zn, -1.0, 1.0)


def _unnorm(c, size):
    return ((c + 1.0) / 2.0) * float(size - 1)


def sample3(field, xn, yn, zn):
    """Trilinear ``grid_sample(align_corners=True)`` of ``field`` [C, Z, Y, X] at
    normalised coords; returns [C, N]. Coordinates are already clamped to [-1, 1]
    so zero padding only ever meets zero weights."""

    C, Z, Y, X = field.shape
    ix = _unnorm(xn, X)
    iy = _unnorm(yn, Y)
    iz = _unnorm(zn, Z)
    x0 = np.floor(ix)
    y0 = np.floor(iy)
    z0 = np.floor(iz)
    fx = ix - x0
    fy = iy - y0
    fz = iz - z0
    x0 = x0.astype(np.int64)
    y0 = y0.astype(np.int64)
    z0 = z0.astype(np.int64)
    x1 = np.minimum(x0 + 1, X - 1)
    y1 = np.minimum(y0 + 1, Y - 1)
    z1 = np.minimum(z0 + 1, Z - 1)
    gx = 1.0 - fx
    gy = 1.0 - fy
    gz = 1.0 - fz
    out = (
        field[:, z0, y0, x0] * (gx * gy * gz)
        + field[:, z0, y0, x1] * (fx * gy * gz)
        + field[:, z0, y1, x0] * (gx * fy * gz)
        + field[:, z0, y1, x1] * (fx * fy * gz)
        + field[:, z1, y0, x0] * (gx * gy * fz)
        + field[:, z1, y0, x1] * (fx * gy * fz)
        + field[:, z1, y1, x0] * (gx * fy * fz)
        + field[:, z1, y1, x1] * (fx * fy * fz)
    )
    return out


def sample2(field, lon, lat, w: Window):
    """Reference ``HannaScheme._bilinear_at`` (hanna.py:1070-1085): bilinear,
    clamp-to-edge, of a [Y, X] field."""

    Y, X = field.shape
    xn = np.clip(2.0 * (lon - w.lon0) / (w.lon1 - w.lon0) - 1.0, -1.0, 1.0)
    yn = np.clip(2.0 * (lat - w.lat0) / (w.lat1 - w.lat0) - 1.0, -1.0, 1.0)
    ix = _unnorm(xn, X)
    iy = _unnorm(yn, Y)
    x0 = np.floor(ix)
    y0 = np.floor(iy)
    fx = ix - x0
    fy = iy - y0
    x0 = x0.astype(np.int64)
    y0 = y0.astype(np.int64)
    x1 = np.minimum(x0 + 1, X - 1)
    y1 = np.minimum(y0 + 1, Y - 1)
    return (
        field[y0, x0] * ((1.0 - fx) * (1.0 - fy))
        + field[y0, x1] * (fx * (1.0 - fy))
        + field[y1, x0] * ((1.0 - fx) * fy)
        + field[y1, x1] * (fx * fy)
    )


def sample_stack(stack, lon, lat, z, w: Window):
    """Reference ``HannaScheme._interp_3d_field`` (hanna.py:1871-1895)."""

    xn, yn, zn = normalize(lon, lat, z, w)
    return sample3(stack, xn, yn, zn)


def m_per_deg_lon(lat):
    """111320 * max(|cos(lat)|, 0.05) (gpu_engine.py:467, hanna.py:884)."""

    return M_PER_DEG_LON_EQ * np.maximum(np.abs(np.cos(np.deg2rad(lat))), 0.05)


# ---------------------------------------------------------------------------
# Hanna free functions (reference hanna.py:142-426)
# ---------------------------------------------------------------------------


def coriolis_parameter(lat):
    return 2.0 * OMEGA_EARTH * np.sin(np.deg2rad(lat))


def air_density(sp, t):
    return sp / (R_DRY * t)


def obukhov_length(ustar, shf, t, sp):
    """hanna.py:154-173; +inf where |kappa g H| <= 1e-10."""

    rho = air_density(sp, t)
    denom = KARMAN * GRAVITY * shf
    num = -(np.maximum(ustar, USTAR_MIN) ** 3) * t * rho * C_P
    finite = np.abs(denom) > 1e-10
    safe = np.where(finite, denom, 1.0)
    return np.where(finite, num / safe, np.inf).astype(ustar.dtype)


def convective_velocity(blh, shf, t, sp):
    """hanna.py:176-189."""

    rho = air_density(sp, t)
    cube = (GRAVITY * np.maximum(blh, BLH_MIN) * shf) / (t * rho * C_P)
    return np.where(cube > 0, np.cbrt(np.maximum(cube, 0.0)), 0.0).astype(blh.dtype)


def _stable(z, blh, ustar):
    """hanna.py:192-222."""

    zoh = np.clip(z / blh, 0.0, 1.0)
    om = np.maximum(1.0 - zoh, 0.0)
    su = 2.0 * ustar * om
    sw = 1.3 * ustar * om
    sv = sw
    zf = np.maximum(zoh, 1e-6)
    tu = 0.15 * blh / np.maximum(su, SIGMA_MIN) * np.sqrt(zf)
    tv = 0.07 * blh / np.maximum(sv, SIGMA_MIN) * np.sqrt(zf)
    tw = 0.10 * blh / np.maximum(sw, SIGMA_MIN) * zf**0.8
    return su, sv, sw, tu, tv, tw


def _neutral(z, ustar, fcor):
    """hanna.py:225-252."""

    us = np.maximum(ustar, USTAR_MIN)
    fa = np.maximum(np.abs(fcor), 1e-5)
    su = 2.0 * ustar * np.exp(-3.0 * fa * z / us)
    sw = 1.3 * ustar * np.exp(-2.0 * fa * z / us)
    sv = sw
    den = 1.0 + 15.0 * fa * z / us
    tw = 0.5 * z / np.maximum(sw, SIGMA_MIN) / np.maximum(den, 1.0)
    return su, sv, sw, tw, tw, tw


def _unstable(z, blh, ustar, wstar, hol):
    """hanna.py:255-298."""

    zoh = np.clip(z / blh, 0.0, 1.0)
    hc = np.clip(hol, -1000.0, 0.0)
    su = np.cbrt(np.maximum(12.0 - 0.5 * hc, 1e-3)) * ustar
    sv = su
    sw2 = 1.2 * wstar**2 * np.maximum(1.0 - 0.9 * zoh, 0.0) * zoh ** (2.0 / 3.0) + (
        1.8 - 1.4 * zoh
    ) * ustar**2
    sw = np.sqrt(np.maximum(sw2, SIGMA_MIN**2))
    tu = 0.15 * blh / np.maximum(su, SIGMA_MIN)
    tv = tu
    azl = np.maximum(zoh * (-hc), 0.0)
    swc = np.maximum(sw, SIGMA_MIN)
    t_near = 0.1 * z / (swc * np.maximum(0.55 - 0.38 * azl, 1e-2))
    t_shallow = 0.59 * z / swc
    t_bulk = 0.15 * blh / swc * (1.0 - np.exp(-5.0 * zoh))
    tw = np.where(azl < 1.0, t_near, np.where(zoh < 0.1, t_shallow, t_bulk))
    return su, sv, sw, tu, tv, tw


def in_bl_sigma_TL(z, blh, ustar, wstar, hol, lat):
    """hanna.py:301-337: all three regimes, selected by h/L (>1 stable, <-1 unstable)."""

    f = coriolis_parameter(lat)
    st = _stable(z, blh, ustar)
    nu = _neutral(z, ustar, f)
    un = _unstable(z, blh, ustar, wstar, hol)
    s_st = hol > 1.0
    s_un = hol < -1.0
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        return tuple(np.where(s_st, a, np.where(s_un, c, b)) for a, b, c in zip(st, nu, un))


def surface_layer_sigma_w(z, ustar, L):
    """hanna.py:405-426."""

    eps = 1e-6
    finite = np.abs(L) > eps
    Ls = np.where(finite, L, 1.0)
    zol = np.where(finite, np.clip(z / Ls, -50.0, 50.0), 0.0)
    s_un = 1.3 * ustar * np.cbrt(np.maximum(1.0 - 3.0 * zol, eps))
    return np.where(zol < 0.0, s_un, 1.3 * ustar)


def column_turbulence(z, blh, ustar, wstar, hol, L, lat, lon, win: Window, p: HannaParams):
    """hanna.py:1500-1596: in-BL profile, optional surface-layer override,
    FLEXPART T_L floors, free-troposphere override for z > h, final floors."""

    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        su, sv, sw, tu, tv, tw = in_bl_sigma_TL(z, blh, ustar, wstar, hol, lat)
        if p.surface_layer_override:
            z_sl = SURFACE_LAYER_FRACTION * blh
            in_sl = z < z_sl
            sw_sl = surface_layer_sigma_w(z, ustar, L)
            zc = np.maximum(z, z_sl)
            su_c, sv_c, *_ = in_bl_sigma_TL(zc, blh, ustar, wstar, hol, lat)
            su = np.where(in_sl, su_c, su)
            sv = np.where(in_sl, sv_c, sv)
            sw = np.where(in_sl, sw_sl, sw)
            zt = np.maximum(z, Z_MIN)
            tu = np.where(in_sl, KARMAN * zt / np.maximum(su, SIGMA_MIN), tu)
            tv = np.where(in_sl, KARMAN * zt / np.maximum(sv, SIGMA_MIN), tv)
            tw = np.where(in_sl, KARMAN * zt / np.maximum(sw, SIGMA_MIN), tw)
    if p.flexpart_tl_floors:
        tu = np.maximum(tu, T_LU_FLOOR)
        tv = np.maximum(tv, T_LV_FLOOR)
        tw = np.maximum(tw, T_LW_FLOOR)
    above = (z > blh) | bool(p.prescribed)
    ft = sample_stack(win.ft, lon, lat, z, win)
    su = np.where(above, ft[0], su)
    sv = np.where(above, ft[0], sv)
    sw = np.where(above, ft[0], sw)
    tu = np.where(above, ft[1], tu)
    tv = np.where(above, ft[1], tv)
    tw = np.where(above, ft[1], tw)
    return (
        np.maximum(su, SIGMA_MIN),
        np.maximum(sv, SIGMA_MIN),
        np.maximum(sw, SIGMA_MIN),
        np.maximum(tu, T_L_MIN),
        np.maximum(tv, T_L_MIN),
        np.maximum(tw, T_L_MIN),
    )


# ---------------------------------------------------------------------------
# Engine primitives (reference gpu_engine.py:26-42, 345-507)
# ---------------------------------------------------------------------------


def ou_update(w_prev, tl, sigma2, dt, eta, drift_dt):
    """gpu_engine.py:26-42: a = exp(-dt/T_L); var = max(sigma^2 (1 - a^2), 0)."""

    a = np.exp(-dt / tl)
    var = np.maximum(sigma2 * (1.0 - a * a), 0.0)
    return a * w_prev + drift_dt + np.sqrt(var) * eta


# ---------------------------------------------------------------------------
# The step (reference hanna.py:855-1068 static path)
# ---------------------------------------------------------------------------


@dataclass
class StepDiag:
    k: np.ndarray  # per-particle substep count (active particles)


def advect_rk2(lon, lat, z, win: Window, alpha: float, dt: float):
    """hanna.py:855-896: x* = x - dt/2 v(x); x_new = x - dt v(x*)."""

    def wind(lo, la, zz):
        xn, yn, zn = normalize(lo, la, zz, win)
        vs = sample3(win.uvw_s, xn, yn, zn)
        ve = sample3(win.uvw_e, xn, yn, zn)
        v = vs * (1.0 - alpha) + ve * alpha
        return v[0] / m_per_deg_lon(la), v[1] / M_PER_DEG_LAT, v[2]

    d0, d1, d2 = wind(lon, lat, z)
    h = 0.5 * dt
    d0b, d1b, d2b = wind(lon - h * d0, lat - h * d1, z - h * d2)
    return lon - dt * d0b, lat - dt * d1b, z - dt * d2b


def step(
    P: dict,
    active: np.ndarray,
    win: Window,
    alpha: float,
    dt: float,
    p: HannaParams,
    *,
    seed: int,
    step_index: int,
    return_diag: bool = False,
):
    """One backward step of the plain method on the particles in ``P`` (a dict
    of [N] arrays: lon, lat, z, up, vp, wp, um, vm, uid). Only ``active``
    particles change (reference gate, hanna.py:1056-1068). Returns a new dict."""

    out = {k: v.copy() for k, v in P.items()}
    idx = np.nonzero(active)[0]
    if idx.size == 0:
        return (out, StepDiag(np.zeros(0, np.int64))) if return_diag else out
    dty = P["lon"].dtype.type
    lon0, lat0, z0 = P["lon"][idx], P["lat"][idx], P["z"][idx]
    uid = P["uid"][idx]

    lon, lat, z = advect_rk2(lon0, lat0, z0, win, alpha, dt)

    # Surface fields, lerped on the grid then sampled (hanna.py:805-813, 955-963).
    surf = [torch_lerp(win.surf_s[c], win.surf_e[c], alpha) for c in range(5)]
    blh = np.maximum(sample2(surf[0], lon, lat, win), BLH_MIN)
    sp = sample2(surf[1], lon, lat, win)
    ustar = np.maximum(sample2(surf[2], lon, lat, win), USTAR_MIN)
    shf = sample2(surf[3], lon, lat, win)
    tsf = sample2(surf[4], lon, lat, win)

    L = obukhov_length(ustar, shf, tsf, sp)
    hol = blh / L
    wstar = convective_velocity(blh, shf, tsf, sp)

    zu = dty(p.z_ubl_m)
    z_eval = np.maximum(z, zu)
    col = (blh, ustar, wstar, hol, L, lat, lon, win, p)
    su, sv, sw, tu, tv, tw = column_turbulence(z_eval, *col)
    z_hi = z_eval + DRIFT_FD_DELTA_M
    z_lo = np.maximum(z_eval - DRIFT_FD_DELTA_M, zu)
    sw_hi = column_turbulence(z_hi, *col)[2]
    sw_lo = column_turbulence(z_lo, *col)[2]
    dsig2 = (sw_hi**2 - sw_lo**2) / np.maximum(z_hi - z_lo, DRIFT_FD_DELTA_M)
    sw_sq = np.maximum(sw**2, SIGMA_MIN**2)

    dens = sample_stack(win.dens, lon, lat, z_eval, win)
    rho = np.maximum(dens[0], 1e-3)
    dens_fwd = sw_sq * dens[1] / rho
    half_b = -0.5 * dsig2
    dens_b = -dens_fwd

    # Substep schedule (hanna.py:1162-1186).
    kreq = np.clip(np.ceil(dt / (p.substep_c * np.maximum(tw, 1e-3))), 1, p.max_substeps).astype(
        np.int64
    )
    sub_dt_full = dty(dt) / kreq.astype(lon.dtype)

    up, vp, wp = P["up"][idx], P["vp"][idx], P["wp"][idx]
    su2, sv2, sw2 = su**2, sv**2, sw**2
    umax, vmax, wmax = W_PRIME_SIGMA_RATIO_MAX * su, W_PRIME_SIGMA_RATIO_MAX * sv, (
        W_PRIME_SIGMA_RATIO_MAX * sw
    )
    x_lon, x_lat, x_z = lon.copy(), lat.copy(), z.copy()
    zero = np.zeros_like(sub_dt_full)
    for i in range(int(kreq.max())):
        owing = kreq > i
        sdt = np.where(owing, sub_dt_full, zero)
        eta = philox.normals(seed, uid, step_index, i).astype(lon.dtype)
        factor = 1.0 + wp**2 / sw_sq
        drift_w = factor * half_b + dens_b
        u_new = ou_update(up, tu, su2, sdt, eta[:, 0], 0.0 * sdt)
        v_new = ou_update(vp, tv, sv2, sdt, eta[:, 1], 0.0 * sdt)
        w_new = ou_update(wp, tw, sw2, sdt, eta[:, 2], drift_w * sdt)
        w_new = np.clip(w_new, -wmax, wmax)
        u_new = np.clip(u_new, -umax, umax)
        v_new = np.clip(v_new, -vmax, vmax)
        x_z = x_z + (-1.0 * w_new) * sdt
        mlon = m_per_deg_lon(x_lat)
        x_lon = x_lon + (-1.0 * u_new) * sdt / mlon
        x_lat = x_lat + (-1.0 * v_new) * sdt / M_PER_DEG_LAT
        below = x_z < 0.0
        x_z = np.where(below, 0.0 - x_z, x_z)
        w_new = np.where(below, -w_new, w_new)
        up, vp, wp = u_new, v_new, w_new

    if p.meander_enabled:
        sm = sample_stack(win.meander, lon, lat, z_eval, win)
        eta = philox.normals(seed, uid, step_index, philox.STREAM_MEANDER).astype(lon.dtype)
        tm = dty(p.meander_timescale_s)
        um = ou_update(P["um"][idx], tm, sm[0] ** 2, dty(dt), eta[:, 0], dty(0.0))
        vm = ou_update(P["vm"][idx], tm, sm[1] ** 2, dty(dt), eta[:, 1], dty(0.0))
        mlon = m_per_deg_lon(x_lat)
        x_lon = x_lon + (-1.0 * um) * dty(dt) / mlon
        x_lat = x_lat + (-1.0 * vm) * dty(dt) / M_PER_DEG_LAT
        out["um"][idx] = um
        out["vm"][idx] = vm

    out["lon"][idx] = x_lon
    out["lat"][idx] = x_lat
    out["z"][idx] = x_z
    out["up"][idx] = up
    out["vp"][idx] = vp
    out["wp"][idx] = wp
    if return_diag:
        return out, StepDiag(kreq)
    return out
