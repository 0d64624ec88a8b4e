"""Numpy restatement of the reduced Emanuel convection (reference
src/lpdm/convection/emanuel.py), plus its per-column application.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference lifts one bbox-mean column; the per-column variant applies the
same matrix construction (``compute_mass_flux_matrix``) and the same backward
sampling (``_redistribute_particles``) to each particle's nearest met column,
with cloud base/top and cloud-base flux given per column.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import philox

GRAVITY = 9.80665
R_DRY = 287.05
R_VAPOR = 461.5
EPS = R_DRY / R_VAPOR
C_P = 1005.0
L_V = 2.5e6
P0 = 100000.0
KAPPA = R_DRY / C_P
P_MIN = 1.0
T_MIN = 150.0
CFL_MAX_OUTFLUX_FRAC = 0.9


def saturation_vapor_pressure(t):
    """Bolton (1980) Eq. 10 (emanuel.py:124-132)."""

    tc = t - 273.15
    return 611.2 * np.exp(17.67 * tc / (tc + 243.5))


def q_to_e(q, p):
    return q * p / (EPS + (1.0 - EPS) * q)


def e_to_q(e, p):
    return EPS * e / (np.maximum(p, P_MIN) - (1.0 - EPS) * e)


def virtual_temperature(t, q):
    return t * (1.0 + (1.0 / EPS - 1.0) * q)


def lcl_temperature_bolton(t, q, p):
    """Bolton (1980) Eq. 22 (emanuel.py:157-170)."""

    e_hpa = np.maximum(q_to_e(q, p), 1.0) / 100.0
    den = 3.5 * np.log(t) - np.log(e_hpa) - 4.805
    return 2840.0 / np.maximum(den, 1e-3) + 55.0


def lcl_pressure_poisson(t, t_lcl, p):
    return p * (t_lcl / np.maximum(t, T_MIN)) ** (C_P / R_DRY)


def potential_temperature(t, p):
    return t * (P0 / np.maximum(p, P_MIN)) ** KAPPA


def equivalent_potential_temperature(t, q, p):
    t_lcl = lcl_temperature_bolton(t, q, p)
    return potential_temperature(t, p) * np.exp(L_V * q / (C_P * np.maximum(t_lcl, T_MIN)))


def saturation_specific_humidity(t, p):
    return e_to_q(saturation_vapor_pressure(t), p)


def lift_parcel_moist(theta_e, p, max_iter=30, t_min=180.0, t_max=350.0):
    """Fixed 30-step bisection (emanuel.py:215-255)."""

    lo = np.full_like(p, t_min)
    hi = np.full_like(p, t_max)
    for _ in range(max_iter):
        mid = 0.5 * (lo + hi)
        with np.errstate(over="ignore"):  # q_sat blows up near p = e_sat: theta_e = inf is "too high"
            th = equivalent_potential_temperature(mid, saturation_specific_humidity(mid, p), p)
        too_high = th > theta_e
        hi = np.where(too_high, mid, hi)
        lo = np.where(too_high, lo, mid)
    return 0.5 * (lo + hi)


def layer_masses(p):
    """|dp|/g with half layers at the ends (emanuel.py:263-279)."""

    p = np.asarray(p)
    dp = np.zeros_like(p)
    dp[1:-1] = 0.5 * np.abs(p[:-2] - p[2:])
    dp[0] = np.abs(p[0] - p[1]) * 0.5
    dp[-1] = np.abs(p[-2] - p[-1]) * 0.5
    return dp / GRAVITY


def compute_lcl_lnb_cape(t, q, p):
    """emanuel.py:282-399 for one ascending column; returns
    (i_lcl, i_lnb, cape, dtv at LCL+1)."""

    n = t.size
    t0, q0, p0 = t[0], q[0], p[0]
    t_lcl = lcl_temperature_bolton(t0, q0, p0)
    p_lcl = lcl_pressure_poisson(t0, t_lcl, p0)
    the = equivalent_potential_temperature(t0, q0, p0)
    th0 = potential_temperature(t0, p0)
    dry = th0 * (p / P0) ** KAPPA
    moist = lift_parcel_moist(np.full_like(p, the), p)
    tpar = np.where(p >= p_lcl, dry, moist)
    tv_env = virtual_temperature(t, q)
    qpar = np.where(p >= p_lcl, q0, saturation_specific_humidity(tpar, p))
    tv_par = virtual_temperature(tpar, qpar)
    buoy = (tv_par - tv_env) / np.maximum(tv_env, T_MIN)
    above = np.nonzero(~(p >= p_lcl))[0]
    if above.size == 0:
        return -1, -1, 0.0, -1.0
    i_lcl = int(above[0])
    dtv = float(tv_par[i_lcl + 1] - tv_env[i_lcl + 1]) if i_lcl + 1 < n else -1.0
    dz = np.zeros(n)
    dz[1:] = R_DRY * 0.5 * (t[:-1] + t[1:]) * (
        np.log(np.maximum(p[:-1], P_MIN)) - np.log(np.maximum(p[1:], P_MIN))) / GRAVITY
    pos = np.maximum(buoy, 0.0)
    pos[:i_lcl] = 0.0
    cape = float((GRAVITY * pos * dz).sum())
    neg = np.nonzero((buoy <= 0.0) & (np.arange(n) > i_lcl))[0]
    i_lnb = int(neg[0]) if neg.size else n - 1
    return i_lcl, i_lnb, cape, dtv


def column_cape(t_s, t_e, q_s, q_e, alpha, p_hpa):
    """compute_lcl_lnb_cape applied to every met column (the reference lifts
    the bbox-mean column, emanuel.py:593-627): t and q lerped at alpha in
    float64 from the [Z, Y, X] float32 fields, q clamped >= Q_MIN_KG_KG = 0
    (emanuel.py:108, 611), pressure p_hpa [Z] * 100. Returns i_lcl, i_lnb
    (int32), cape, dtv (float64), each [Y, X]."""

    al = float(alpha)
    t = np.asarray(t_s, np.float64) * (1.0 - al) + np.asarray(t_e, np.float64) * al
    q = np.maximum(np.asarray(q_s, np.float64) * (1.0 - al) + np.asarray(q_e, np.float64) * al, 0.0)
    p = np.asarray(p_hpa, np.float64) * 100.0
    ny, nx = t.shape[1:]
    out = [np.zeros((ny, nx), np.int32), np.zeros((ny, nx), np.int32), np.zeros((ny, nx)), np.zeros((ny, nx))]
    for j in range(ny):
        for i in range(nx):
            for arr, v in zip(out, compute_lcl_lnb_cape(t[:, j, i], q[:, j, i], p)):
                arr[j, i] = v
    return tuple(out)


def conv_descriptors_from_lift(i_lcl, i_lnb, cape, dtv, levels):
    """The [4, Y, X] float32 (CAPE, LCL height, top height, T_v excess) fields
    column_descriptors takes, from a parcel lift; heights +inf for -1."""

    lev = np.asarray(levels, np.float32)
    inf = np.float32(np.inf)
    lcl_h = np.where(i_lcl >= 0, lev[np.maximum(i_lcl, 0)], inf)
    top_h = np.where(i_lnb >= 0, lev[np.maximum(i_lnb, 0)], inf)
    return np.stack([np.asarray(cape, np.float32), lcl_h, top_h, np.asarray(dtv, np.float32)]).astype(np.float32)


def detrainment_weights(i_lcl, i_lnb, n):
    w = np.zeros(n)
    depth = i_lnb - i_lcl
    if depth <= 0:
        return w
    for k in range(i_lcl, i_lnb + 1):
        w[k] = 1.0 - (k - i_lcl) / depth
    return w


def mass_flux_matrix(m, i_lcl, i_lnb, m_b):
    """emanuel.py:429-506 given the layer masses ``m``; m_b already includes
    the event interval (kg/m^2)."""

    n = m.size
    f = np.zeros((n, n))
    if i_lcl < 0 or i_lnb <= i_lcl or m_b <= 0.0:
        return f
    e = np.zeros(n)
    e[:i_lcl] = m_b * m[:i_lcl] / max(m[:i_lcl].sum(), 1e-12)
    w = detrainment_weights(i_lcl, i_lnb, n)
    d = m_b * w / max(w.sum(), 1e-12)
    f = f + np.outer(e, d) / m_b
    phi = np.maximum((np.cumsum(e) - np.cumsum(d))[:-1], 0.0)
    f = f + np.diag(phi, -1)
    mx = (f.sum(axis=1) / np.maximum(m, 1e-12)).max()
    if mx > CFL_MAX_OUTFLUX_FRAC:
        f = f * (CFL_MAX_OUTFLUX_FRAC / mx)
    return f


def mass_flux_matrix_f32(m, i_lcl, i_lnb, m_b):
    """mass_flux_matrix (emanuel.py:429-506) in float32 arithmetic, as the
    reference computes it on float32 met: every term rounded to float32, sums
    taken sequentially in level order. Returns the float32 entries."""

    f32 = np.float32
    m = np.asarray(m, dtype=np.float32)
    n = m.size
    out = np.zeros((n, n), dtype=np.float32)
    mb = f32(m_b)
    if i_lcl < 0 or i_lnb <= i_lcl or mb <= 0:
        return out
    bl = f32(0)
    for i in range(i_lcl):
        bl = f32(bl + m[i])
    bl = max(bl, f32(1e-12))
    depth = f32(i_lnb - i_lcl)
    w = np.zeros(n, dtype=np.float32)
    for k in range(i_lcl, i_lnb + 1):
        w[k] = f32(1) - f32(k - i_lcl) / depth
    wt = f32(0)
    for k in range(i_lcl, i_lnb + 1):
        wt = f32(wt + w[k])
    wt = max(wt, f32(1e-12))
    e = np.zeros(n, dtype=np.float32)
    for i in range(i_lcl):
        e[i] = f32(mb * m[i]) / bl
    d = np.array([f32(mb * w[k]) / wt for k in range(n)], dtype=np.float32)
    sum_d = f32(0)
    for k in range(n):
        sum_d = f32(sum_d + d[k])
    phi = np.zeros(n, dtype=np.float32)
    ce = cd = f32(0)
    mx = f32(0)
    prev = f32(0)
    for i in range(n):
        row = f32(f32(e[i] * sum_d) / mb) + prev
        mx = max(mx, f32(row / max(m[i], f32(1e-12))))
        ce = f32(ce + e[i])
        cd = f32(cd + d[i])
        prev = max(f32(ce - cd), f32(0)) if i < n - 1 else f32(0)
        phi[i] = prev
    scale = f32(CFL_MAX_OUTFLUX_FRAC) / mx if mx > f32(CFL_MAX_OUTFLUX_FRAC) else f32(1)
    for i in range(n):
        for j in range(n):
            v = f32(f32(e[i] * d[j]) / mb)
            if i == j + 1:
                v = f32(v + phi[j])
            out[i, j] = f32(scale * v)
    return out


def column_sum_error(f, m):
    """max_i |sum_j f_ij - sum_j f_ji| / m_i, summed in float64 in index order
    (the per-column non-divergence check, emanuel.py:437-446)."""

    f = np.asarray(f, dtype=np.float64)
    n = f.shape[0]
    worst = 0.0
    for i in range(n):
        row = col = 0.0
        for j in range(n):
            row += f[i, j]
            col += f[j, i]
        worst = max(worst, abs(row - col) / max(float(m[i]), 1e-12))
    return worst


def backward_probabilities(f, layer_mass_clamped):
    """Transposed, row-normalised transition (emanuel.py:756-763)."""

    lm = layer_mass_clamped
    ff = f + np.diag(np.maximum(lm - f.sum(axis=1), 0.0))
    return np.maximum(ff.T / np.maximum(lm, 1e-12)[:, None], 0.0)


def sample_destinations(prob, host, u):
    """emanuel.py:778-787: renormalise row ``host``, cumsum, searchsorted(u)."""

    rows = prob[host]
    rows = rows / np.maximum(rows.sum(axis=1, keepdims=True), 1e-12)
    cum = np.cumsum(rows, axis=1)
    dest = (cum < u[:, None]).sum(axis=1)
    return np.minimum(dest, prob.shape[0] - 1)


def host_levels(z, height):
    """Nearest level (emanuel.py:765-771)."""

    n = height.size
    up = np.clip(np.searchsorted(height, z, side="left"), 1, n - 1)
    lo = up - 1
    return np.where((z - height[lo]) < (height[up] - z), lo, up)


def column_descriptors(conv, t_s, t_e, alpha, p_hpa, levels, closure_c=0.03, trigger=0.9,
                       min_cape=50.0, min_depth=500.0):
    """Per-column (i_lcl, i_lnb, m_b) from CAPE / LCL / cloud-top / T_v-excess
    fields, with the reference's trigger and closure (emanuel.py:637-675).
    Computed in float32 like the rest of the engine's per-column setup."""

    f32 = np.float32
    cape, lcl_h, top_h, dtv = (np.asarray(c, f32) for c in conv)
    lev = np.asarray(levels, f32)
    Z = lev.size
    ge_l = lev[:, None, None] >= lcl_h[None]
    ge_t = lev[:, None, None] >= top_h[None]
    lcl = np.where(ge_l.any(0), ge_l.argmax(0), -1)
    lnb = np.where(ge_t.any(0), ge_t.argmax(0), Z - 1)
    fires = (lcl >= 0) & (lnb > lcl) & (cape >= f32(min_cape)) & (dtv >= f32(trigger))
    lcl_c = np.maximum(lcl, 0)
    depth = lev[lnb] - lev[lcl_c]
    fires &= ~(depth < f32(min_depth))
    yy, xx = np.indices(lcl.shape)
    al = f32(alpha)
    tl = t_s[lcl_c, yy, xx].astype(f32) * (f32(1.0) - al) + t_e[lcl_c, yy, xx].astype(f32) * al
    p_l = np.asarray(p_hpa, np.float64)[lcl_c].astype(f32) * f32(100.0)
    rho = p_l / (f32(287.05) * np.maximum(tl, f32(150.0)))
    wb = np.minimum(np.sqrt(f32(2.0) * np.maximum(cape, f32(0.0))), f32(5.0))
    mb = f32(closure_c) * rho * wb * f32(3600.0)
    return (np.where(fires, lcl, -1), np.where(fires, lnb, -1), np.where(fires, mb, f32(0.0)))


def nearest_column(lon, lat, lon0, lon1, lat0, lat1, nx, ny):
    """Nearest met column index of float32 positions (clamp-to-edge)."""

    f32 = np.float32
    sx = f32((nx - 1) / (lon1 - lon0))
    sy = f32((ny - 1) / (lat1 - lat0))
    ix = np.clip((np.asarray(lon, f32) - f32(lon0)) * sx, f32(0), f32(nx - 1))
    iy = np.clip((np.asarray(lat, f32) - f32(lat0)) * sy, f32(0), f32(ny - 1))
    cx = np.minimum(np.floor(ix + f32(0.5)).astype(np.int64), nx - 1)
    cy = np.minimum(np.floor(iy + f32(0.5)).astype(np.int64), ny - 1)
    return cy * nx + cx


def convect_columns(z, active, col, i_lcl, i_lnb, m_b, p_hpa, levels, seed, uid, step_index,
                    dtype=np.float64):
    """Backward redistribution of ``z`` for active particles, each through its
    column's matrix. Returns the new z array."""

    lm = layer_masses(np.asarray(p_hpa, dtype) * 100.0)
    lmc = np.maximum(lm, 1e-6)
    h = np.asarray(levels, dtype)
    z = np.asarray(z, dtype).copy()
    ids = np.nonzero(active)[0]
    if ids.size == 0:
        return z
    u = philox.uniforms(seed, np.asarray(uid)[ids], step_index, philox.STREAM_CONVECTION)
    cols = np.asarray(col)[ids]
    lcl = np.asarray(i_lcl).reshape(-1)[cols]
    lnb = np.asarray(i_lnb).reshape(-1)[cols]
    mb = np.asarray(m_b).reshape(-1)[cols]
    fire = (lcl >= 0) & (lnb > lcl) & (mb > 0)
    groups = {}
    for j in np.nonzero(fire)[0]:
        groups.setdefault((int(lcl[j]), int(lnb[j]), float(mb[j])), []).append(j)
    for (a, b, mbv), js in groups.items():
        js = np.asarray(js)
        f = mass_flux_matrix(lm, a, b, mbv)
        prob = backward_probabilities(f, lmc)
        zz = z[ids[js]]
        host = host_levels(zz, h)
        dest = sample_destinations(prob, host, u[js, 0])
        moved = dest != host
        zl = h[np.maximum(dest - 1, 0)]
        zh = h[dest]
        znew = zl + u[js, 1] * (zh - zl)
        z[ids[js]] = np.where(moved, znew, zz)
    return z


def maybe_convect_bbox(t_col, q_col, p_pa, height_col, closure_c=0.03, trigger=0.9,
                       min_cape=50.0, min_depth=500.0):
    """Reference bbox-mean trigger and closure (emanuel.py:629-675): returns
    (fires, i_lcl, i_lnb, m_b * 3600) for an ascending column."""

    i_lcl, i_lnb, cape, dtv = compute_lcl_lnb_cape(t_col, q_col, p_pa)
    fires = i_lcl >= 0 and i_lnb > i_lcl and cape >= min_cape and dtv >= trigger
    if fires and float(height_col[i_lnb] - height_col[i_lcl]) < min_depth:
        fires = False
    if not fires:
        return False, i_lcl, i_lnb, 0.0
    rho = p_pa[i_lcl] / (R_DRY * max(t_col[i_lcl], T_MIN))
    wb = min(math.sqrt(2.0 * max(cape, 0.0)), 5.0)
    return True, i_lcl, i_lnb, closure_c * rho * wb * 3600.0
