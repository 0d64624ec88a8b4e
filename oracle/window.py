"""Numpy restatement of the per-window support stacks built at the window
midpoint: free-troposphere sigma/T_L, density and its vertical gradient, and the
meander sigma.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""

from __future__ import annotations

import numpy as np

from oracle.lpdm import GRAVITY, KARMAN, R_DRY, C_P, SIGMA_MIN, T_L_MIN

P0 = 100000.0
KAPPA_POISSON = R_DRY / C_P
FT_LAMBDA = 100.0
FT_RI_CRIT = 0.25
FT_KZ_FLOOR = 0.1
FT_KZ_CEIL = 50.0
FT_SHEAR_FLOOR = 1e-8
FT_N2_FLOOR = 1e-6
FT_TL_DEFAULT = 100.0
FT_TL_MAX = 1000.0


def vertical_gradient(f, h):
    """hanna.py:1855-1869: central differences inside, one-sided at the ends;
    denominators with |d| < 1e-6 are replaced by 1e-6."""

    def safe(d):
        return np.where(np.abs(d) < 1e-6, 1e-6, d).astype(f.dtype)

    g = np.empty_like(f)
    g[1:-1] = (f[2:] - f[:-2]) / safe(h[2:] - h[:-2])
    g[0] = (f[1] - f[0]) / safe(h[1] - h[0])
    g[-1] = (f[-1] - f[-2]) / safe(h[-1] - h[-2])
    return g


def potential_temperature(t, p):
    return t * (P0 / p) ** KAPPA_POISSON


def brunt_vaisala_squared(theta, dtheta_dz):
    return (GRAVITY / np.maximum(theta, 1.0)) * dtheta_dz


def gradient_richardson(n2, shear_sq):
    return n2 / np.maximum(shear_sq, FT_SHEAR_FLOOR)


def free_trop_diffusivity(z, shear_mag, ri):
    """hanna.py:359-382."""

    ml = (KARMAN * z) / (1.0 + KARMAN * z / FT_LAMBDA)
    f_st = np.maximum(1.0 - ri / FT_RI_CRIT, 0.0) ** 2
    f_un = np.sqrt(np.maximum(1.0 - 16.0 * ri, 0.0))
    f = np.where(ri < 0.0, f_un, f_st)
    return np.clip(ml**2 * shear_mag * f, FT_KZ_FLOOR, FT_KZ_CEIL)


def free_trop_sigma_TL(k_z, n2):
    """hanna.py:385-402."""

    with np.errstate(divide="ignore", invalid="ignore"):
        tl = np.where(
            n2 > FT_N2_FLOOR,
            np.clip(0.5 / np.sqrt(np.maximum(n2, FT_N2_FLOOR)), T_L_MIN, FT_TL_MAX),
            FT_TL_DEFAULT,
        ).astype(k_z.dtype)
    sw = np.sqrt(np.maximum(k_z / tl, SIGMA_MIN**2))
    return sw, tl


def window_mid(start, end):
    """hanna.py:1682-1692."""

    return 0.5 * (start + end)


def free_trop_fields(t_mid, u_mid, v_mid, height, pressure_pa):
    """hanna.py:1694-1745 (``_build``); returns [2, Z, Y, X] = (sigma_w, T_Lw).
    ``pressure_pa`` is [Z] (horizontally uniform per level)."""

    p = pressure_pa[:, None, None]
    theta = potential_temperature(t_mid, p)
    n2 = brunt_vaisala_squared(theta, vertical_gradient(theta, height))
    du = vertical_gradient(u_mid, height)
    dv = vertical_gradient(v_mid, height)
    shear_sq = du**2 + dv**2
    shear_mag = np.sqrt(np.maximum(shear_sq, FT_SHEAR_FLOOR))
    ri = gradient_richardson(n2, shear_sq)
    k_z = free_trop_diffusivity(np.maximum(height, 1.0), shear_mag, ri)
    sw, tl = free_trop_sigma_TL(k_z, n2)
    return np.stack([sw, tl])


def density_fields(t_mid, height, pressure_pa):
    """hanna.py:1747-1784; returns [2, Z, Y, X] = (rho, drho/dz)."""

    p = pressure_pa[:, None, None]
    rho = p / (R_DRY * np.maximum(t_mid, 1.0))
    return np.stack([rho, vertical_gradient(rho, height)])


def windowed_std(field, r: int):
    """hanna.py:1813-1830: population std over the edge-clipped (2r+1)^2 window
    (avg_pool2d with count_include_pad=False), via mean(x^2) - mean(x)^2."""

    Z, Y, X = field.shape
    s1 = np.zeros_like(field)
    s2 = np.zeros_like(field)
    cnt = np.zeros((Y, X), dtype=field.dtype)
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            ys = slice(max(0, dy), Y + min(0, dy))
            yd = slice(max(0, -dy), Y + min(0, -dy))
            xs = slice(max(0, dx), X + min(0, dx))
            xd = slice(max(0, -dx), X + min(0, -dx))
            s1[:, yd, xd] += field[:, ys, xs]
            s2[:, yd, xd] += field[:, ys, xs] ** 2
            cnt[yd, xd] += 1.0
    mean = s1 / cnt
    var = np.maximum(s2 / cnt - mean**2, 0.0)
    return np.sqrt(var)


def meander_fields(u_mid, v_mid, coeff: float = 0.16, r: int = 1):
    """hanna.py:1786-1811; returns [2, Z, Y, X]."""

    return np.stack([coeff * windowed_std(u_mid, r), coeff * windowed_std(v_mid, r)])
