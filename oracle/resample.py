"""Numpy restatement of the hourly terrain-following resample.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows reference ``vertical_grid.py:70-316`` and ``met_reader.py:641-693``.
The 3-D channels are lerped in float32 with float64-derived fractions, the
``w`` slope correction is evaluated in float64 and rounded once to float32, as in
the reference; the HIP column kernel is held to these values bit for bit.
"""

from __future__ import annotations

import numpy as np

M_PER_DEG_LAT = 110540.0
M_PER_DEG_LON_EQ = 111320.0


def stretched_agl_levels(n_levels: int, alt_max_m: float, first_layer_m: float = 10.0):
    """vertical_grid.py:70-144 (ratio by 200-step bisection, top pinned)."""

    n_layers = n_levels - 1
    if n_layers == 1:
        return np.array([0.0, float(alt_max_m)])

    def span(r):
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
    ratio = 0.5 * (lo + hi)
    th = first_layer_m * ratio ** np.arange(n_layers, dtype="float64")
    levels = np.concatenate(([0.0], np.cumsum(th)))
    levels[-1] = float(alt_max_m)
    return levels


def regrid_weights(height_agl_p, agl_levels):
    """vertical_grid.py:166-221. Returns (lower [Za,Y,X] int64 in the ASCENDING
    level axis, frac [Za,Y,X] float64, flipped)."""

    h = np.asarray(height_agl_p, dtype=np.float64)
    Zp, Y, X = h.shape
    mean_profile = np.nanmean(h, axis=(1, 2))
    flipped = bool(mean_profile[0] > mean_profile[-1])
    if flipped:
        h = h[::-1]
    rows = np.ascontiguousarray(h.transpose(1, 2, 0).reshape(-1, Zp))
    tg = np.asarray(agl_levels, dtype=np.float64)
    # searchsorted(right=False) per row == count of levels strictly below target.
    idx = (rows[:, :, None] < tg[None, None, :]).sum(axis=1)  # [YX, Za]
    k0 = (rows < 0.0).sum(axis=1, keepdims=True)
    lower = np.minimum(np.maximum(np.maximum(idx - 1, 0), k0), Zp - 2)
    h_lo = np.take_along_axis(rows, lower, axis=1)
    h_hi = np.take_along_axis(rows, lower + 1, axis=1)
    frac = np.clip((tg[None, :] - h_lo) / np.maximum(h_hi - h_lo, 1e-6), 0.0, 1.0)
    Za = tg.size
    lower = lower.reshape(Y, X, Za).transpose(2, 0, 1)
    frac = frac.reshape(Y, X, Za).transpose(2, 0, 1)
    return np.ascontiguousarray(lower), np.ascontiguousarray(frac), flipped


def lerp_tensor(a, b, w):
    """ATen lerp with a tensor weight, two-sided form, no fused multiply-add."""

    return np.where(np.abs(w) < 0.5, a + w * (b - a), b - (b - a) * (1.0 - w)).astype(a.dtype)


def apply_regrid(field_p, lower, frac, flipped):
    """vertical_grid.py:224-257 for [C, Zp, Y, X]."""

    f = np.asarray(field_p)
    Zp = f.shape[1]
    if flipped:
        lo = (Zp - 1) - lower
        hi = lo - 1
    else:
        lo = lower
        hi = lower + 1
    f_lo = np.take_along_axis(f, np.broadcast_to(lo[None], (f.shape[0],) + lo.shape), axis=1)
    f_hi = np.take_along_axis(f, np.broadcast_to(hi[None], (f.shape[0],) + hi.shape), axis=1)
    w = frac.astype(f.dtype)[None]
    return lerp_tensor(f_lo, f_hi, w)


def m_per_deg_lon_rows(lat_deg):
    """vertical_grid.py:290: clip(111320 cos(lat), 1, None) per row, float64."""

    return np.clip(M_PER_DEG_LON_EQ * np.cos(np.deg2rad(np.asarray(lat_deg, np.float64))), 1.0, None)


def terrain_gradient(terrain_m, lat_deg, lon_deg):
    """vertical_grid.py:280-293 (numpy.gradient with coordinate arrays).

    The reference's terrain is float32 (z_sfc / g), so numpy keeps the slopes in
    float32; dh/dx becomes float64 on division by the float64 row metric while
    dh/dy stays float32 (a Python-float divisor)."""

    dh_dlat, dh_dlon = np.gradient(np.asarray(terrain_m, np.float32), lat_deg, lon_deg)
    return dh_dlon / m_per_deg_lon_rows(lat_deg)[:, None], dh_dlat / M_PER_DEG_LAT


def slope_correct_w(w_agl, u_agl, v_agl, dhdx, dhdy, agl_levels, z_top_m):
    """vertical_grid.py:296-316; float64 arithmetic, result stored as w's dtype."""

    taper = np.clip(1.0 - np.asarray(agl_levels, np.float64) / max(float(z_top_m), 1.0), 0.0, 1.0)
    slope = u_agl * dhdx[None] + v_agl * dhdy[None]
    return (w_agl - taper[:, None, None] * slope).astype(w_agl.dtype)


def resample_hour(src3, src2, height_agl_p, terrain_m, lat, lon, agl_levels, p_level_pa,
                  uvw_index=(0, 1, 2)):
    """met_reader.py:641-693 for one hour.

    src3: [C3, Zp, Y, X] float32 3-D channels (u, v, w, ... in that order);
    src2: [C2, Y, X] float32 2-D channels (surface fields, passed through: the
    reference broadcasts them over Z and the lerp of equal values returns them);
    p_level_pa: [Zp] pressure per source level (Pa, float64).

    Returns (out3 [C3, Za, Y, X] float32, src2, pressure_level_hpa [Za] float64).
    """

    lower, frac, flipped = regrid_weights(height_agl_p, agl_levels)
    out3 = apply_regrid(np.asarray(src3, np.float32), lower, frac, flipped)
    dhdx, dhdy = terrain_gradient(terrain_m, lat, lon)
    iu, iv, iw = uvw_index
    out3[iw] = slope_correct_w(out3[iw], out3[iu], out3[iv], dhdx, dhdy, agl_levels,
                               float(agl_levels[-1]))
    Zp, Y, X = np.asarray(height_agl_p).shape
    p3 = np.broadcast_to((np.asarray(p_level_pa, np.float64) / 100.0)[:, None, None], (Zp, Y, X))
    p_agl = apply_regrid(np.ascontiguousarray(p3)[None], lower, frac, flipped)[0]
    return out3, np.asarray(src2, np.float32), p_agl.mean(axis=(1, 2))


# ---------------------------------------------------------------------------
# Per-hour unit conversions of raw source fields (the reader's host side)
# ---------------------------------------------------------------------------

GRAVITY = 9.80665
R_DRY = 287.05


def agl_from_geopotential(geopotential, geopotential_sfc):
    """met_reader.py:1052-1078: (z - z_sfc) / g on the float32 fields (numpy
    float32 arithmetic, the scalar g taken as float32); not clamped."""

    z = np.asarray(geopotential, np.float32)
    zs = np.asarray(geopotential_sfc, np.float32)
    return (z - zs[None]) / np.float32(GRAVITY)


def omega_to_w(omega, temperature, p_level_pa):
    """met_reader.py:1268-1300: w = -(R_d T / (g p)) omega in float64, p the
    level pressure broadcast to [Z, Y, X]; returned as float32 (the tensor the
    reader builds)."""

    om = np.asarray(omega, np.float64)
    t = np.asarray(temperature, np.float64)
    p3 = np.broadcast_to(np.asarray(p_level_pa, np.float64)[:, None, None], om.shape)
    return (-(R_DRY * t / (GRAVITY * p3)) * om).astype(np.float32)


def shf_to_upward(shf, units: str, accumulation_seconds: float = 3600.0):
    """met_reader.py:1353-1385: de-accumulate J/m^2 over the accumulation
    window, then flip ECMWF's downward-positive sign; float32 arithmetic.
    ``units``: "accumulated" (J/m^2) or "instantaneous" (W/m^2)."""

    a = np.asarray(shf, np.float32)
    if units == "accumulated":
        a = a / np.float32(accumulation_seconds)
    elif units != "instantaneous":
        raise ValueError(f"unknown sensible heat flux units {units!r}")
    return -a


def convert_source_hour(sh, w_ch: int = 2, t_ch: int = 3, shf_ch: int = 3):
    """(src3, src2, height_agl) of a raw source hour (``sh.units == "source"``:
    omega in the w channel, accumulated downward sensible heat flux, geopotential
    instead of heights) in the resample's units; other hours pass through."""

    if getattr(sh, "units", "final") != "source":
        return sh.src3, sh.src2, sh.height_agl
    src3 = np.array(sh.src3, np.float32, copy=True)
    src2 = np.array(sh.src2, np.float32, copy=True)
    src3[w_ch] = omega_to_w(src3[w_ch], src3[t_ch], sh.p_level_pa)
    src2[shf_ch] = shf_to_upward(src2[shf_ch], "accumulated")
    return src3, src2, agl_from_geopotential(sh.geopotential, sh.geopotential_sfc)


def resample_source_hour(sh, met, levels):
    """resample_hour of a source hour in either units (converted first)."""

    src3, src2, height = convert_source_hour(sh)
    return resample_hour(src3, src2, height, met.terrain, met.lat, met.lon, levels, sh.p_level_pa)
