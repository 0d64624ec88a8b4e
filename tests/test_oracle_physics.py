"""Pins the oracle's engine primitives and Hanna functions with the reference's
own known-answer tests (re-expressed against the oracle; same inputs and
tolerances), closed forms and invariants.

Reference tests: tests/test_physics.py, tests/test_hanna.py."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import lpdm as O
from oracle import philox
from oracle import window as W

f64 = np.float64


def a(*v):
    return np.array(v, dtype=f64)


# --- Hanna free functions (reference tests/test_hanna.py:48-364) ------------

def test_coriolis_parameter_signs_and_magnitude():
    f = O.coriolis_parameter(a(0.0, 30.0, 45.0, 60.0, -45.0))
    assert abs(f[0]) < 1e-10
    assert abs(f[2] - 2.0 * 7.2921e-5 * math.sin(math.radians(45.0))) < 1e-12
    assert f[3] > f[2] > f[1] > 0 and f[4] < 0 and abs(f[4] + f[2]) < 1e-12


def test_air_density_standard_conditions():
    assert abs(O.air_density(a(101325.0), a(288.0))[0] - 1.225) < 0.01


def test_obukhov_length_signs():
    L = O.obukhov_length(a(0.4, 0.4, 0.4), a(100.0, -50.0, 0.0), a(288.0, 288, 288), a(101325.0, 101325, 101325))
    assert L[0] < 0 and L[1] > 0 and math.isinf(L[2])


def test_convective_velocity_closed_form():
    w = O.convective_velocity(a(1000.0, 1000, 1000), a(200.0, -50.0, 0.0), a(288.0, 288, 288), a(101325.0, 101325, 101325))
    assert w[0] > 0 and w[1] == 0.0 and w[2] == 0.0
    rho = 101325.0 / (287.05 * 288.0)
    assert abs(w[0] - ((9.80665 * 1000.0 * 200.0) / (288.0 * rho * 1005.0)) ** (1 / 3)) < 1e-6


def test_in_bl_sigma_at_z_zero():
    su, sv, sw, *_ = O.in_bl_sigma_TL(a(0, 0, 0), a(1000.0, 1000, 1000), a(0.5, 0.5, 0.5), a(0, 0, 1.5),
                                      a(5.0, 0.0, -5.0), a(45.0, 45, 45))
    assert abs(su[0] - 1.0) < 1e-9 and abs(sw[0] - 0.65) < 1e-9 and abs(sv[0] - 0.65) < 1e-9
    assert abs(su[1] - 1.0) < 1e-9 and abs(sw[1] - 0.65) < 1e-9 and abs(sv[1] - 0.65) < 1e-9
    assert abs(sw[2] - math.sqrt(1.8 * 0.25)) < 1e-6
    assert not np.allclose(su, sv)


def test_in_bl_stable_sigma_vanishes_at_bl_top():
    sw = O.in_bl_sigma_TL(a(1000.0), a(1000.0), a(0.5), a(0.0), a(5.0), a(45.0))[2]
    assert sw[0] < 1e-2


def test_in_bl_stable_timescales_match_hanna():
    us, h = 0.3, 500.0
    z = a(25.0, 100.0, 250.0, 400.0)
    one = np.ones_like(z)
    su, sv, sw, tu, tv, tw = O.in_bl_sigma_TL(z, h * one, us * one, 0 * one, 5 * one, 45 * one)
    for i, zv in enumerate(z):
        zeta = zv / h
        su_e, sv_e = 2.0 * us * (1 - zeta), 1.3 * us * (1 - zeta)
        assert abs(su[i] - su_e) < 1e-9 and abs(sv[i] - sv_e) < 1e-9 and abs(sw[i] - sv_e) < 1e-9
        assert abs(tu[i] - 0.15 * h / su_e * zeta**0.5) < 1e-6
        assert abs(tv[i] - 0.07 * h / sv_e * zeta**0.5) < 1e-6
        assert abs(tw[i] - 0.10 * h / sv_e * zeta**0.8) < 1e-6
    assert np.allclose(tv / tu, 0.718, rtol=1e-3)


def test_in_bl_regime_selection_at_boundaries():
    def su(hol, ws):
        return O.in_bl_sigma_TL(a(100.0), a(1000.0), a(0.5), a(ws), a(hol), a(45.0))[0][0]
    assert su(-10.0, 1.5) > su(-1.5, 1.5) > 0
    assert su(0.5, 0.0) != su(1.5, 0.0)
    # exactly +-1 is neutral (strict inequalities, hanna.py:323-324)
    assert su(1.0, 0.0) == su(0.0, 0.0) and su(-1.0, 0.0) == su(0.0, 0.0)


def test_surface_layer_sigma_w_regimes():
    sw = O.surface_layer_sigma_w(a(10.0, 10, 10), a(0.5, 0.5, 0.5), a(np.inf, -100.0, 200.0))
    assert abs(sw[0] - 0.65) < 1e-6 and abs(sw[1] - 0.65 * 1.3 ** (1 / 3)) < 1e-6 and abs(sw[2] - 0.65) < 1e-6


def test_free_trop_closure_known_answers():
    kappa = 287.05 / 1005.0
    th = W.potential_temperature(a(280.0, 250.0), a(100000.0, 50000.0))
    assert abs(th[0] - 280.0) < 1e-9 and abs(th[1] - 250.0 * 2.0**kappa) < 1e-6
    assert W.brunt_vaisala_squared(a(300.0), a(0.005))[0] > 0 > W.brunt_vaisala_squared(a(300.0), a(-0.005))[0]
    shear = a(0.01, 0.01, 0.01)
    ri = W.gradient_richardson(a(0.0, 1e-5, 1e-3), shear**2)
    k = W.free_trop_diffusivity(a(1000.0, 1000, 1000), shear, ri)
    assert k[0] > k[1] >= k[2] and abs(k[2] - 0.1) < 1e-9 and k[0] <= 50.0
    sw, tl = W.free_trop_sigma_TL(a(0.1, 1.0, 10.0), a(1e-4, 1e-4, 1e-4))
    assert np.allclose(sw**2 * tl, a(0.1, 1.0, 10.0), rtol=1e-5)
    assert W.free_trop_diffusivity(a(2000.0), a(0.001), a(100.0))[0] >= 0.1
    # N^2 <= floor -> default 100 s timescale
    assert W.free_trop_sigma_TL(a(1.0), a(0.0))[1][0] == 100.0


def test_windowed_std_matches_numpy_population_std():
    rng = np.random.default_rng(0)
    arr = rng.standard_normal((1, 6, 7))
    got = W.windowed_std(arr, 1)
    want = np.empty_like(arr)
    for j in range(6):
        for i in range(7):
            want[0, j, i] = arr[0, max(0, j - 1): j + 2, max(0, i - 1): i + 2].std()
    np.testing.assert_allclose(got, want, atol=1e-10)
    assert np.allclose(W.windowed_std(np.full((3, 5, 5), 7.0), 1), 0.0, atol=1e-6)


def test_vertical_gradient_exact_on_linear_fields():
    h = np.cumsum(np.r_[0.0, np.random.default_rng(2).uniform(5, 500, 9)])[:, None, None] * np.ones((1, 3, 4))
    f = 3.0 + 0.02 * h
    assert np.allclose(W.vertical_gradient(f, h), 0.02, atol=1e-12)


# --- Engine primitives (reference tests/test_physics.py) -------------------

def test_ou_drift_is_deterministic_increment():
    w = a(0.2, -0.1, 0.0)
    tl, s2, dt = 100.0, 1.0, 10.0
    aa = math.exp(-dt / tl)
    assert np.allclose(O.ou_update(w, tl, s2, dt, 0 * w, 0.0), aa * w, atol=1e-12)
    d = a(0.01, 0.02, -0.03)
    assert np.allclose(O.ou_update(w, tl, s2, dt, 0 * w, d * dt), aa * w + d * dt, atol=1e-12)
    # dt = 0 is the identity (the static path's no-op substep)
    eta = a(1.5, -0.3, 2.2)
    assert np.array_equal(O.ou_update(w, tl, s2, 0.0, eta, 0.0 * d), w)


def test_level_lookup_known_answers():
    lev = a(0.0, 100.0, 300.0, 700.0, 1500.0, 3000.0)
    win = O.Window(*(None,) * 7, lon0=-1.0, lon1=1.0, lat0=-1.0, lat1=1.0, levels=lev)
    for z, want in ((300.0, -0.2), (200.0, -0.4), (-50.0, -1.0), (9000.0, 1.0), (0.0, -1.0)):
        zn = O.normalize(a(0.0), a(0.0), a(z), win)[2][0]
        assert abs(zn - want) < 1e-6, (z, zn)


def test_uniform_wind_rk2_exact_and_metric():
    Z, Y, X = 5, 6, 7
    uvw = np.zeros((3, Z, Y, X))
    uvw[0], uvw[1], uvw[2] = 2.5, -1.25, 0.3
    lev = np.linspace(0.0, 2000.0, Z)
    win = O.Window(uvw, uvw, None, None, None, None, None, 0.0, 3.0, 50.0, 48.5, lev)
    lon, lat, z = a(1.0, 2.0), a(49.0, 49.6), a(100.0, 700.0)
    dt = 60.0
    lo, la, zz = O.advect_rk2(lon, lat, z, win, 0.3, dt)
    assert np.allclose(zz, z - 0.3 * dt, atol=1e-9)
    # first stage lat shift, then the metric at the midpoint latitude
    lat_mid = lat - 0.5 * dt * (-1.25 / 110540.0)
    assert np.allclose(la, lat - dt * (-1.25 / 110540.0), atol=1e-12)
    assert np.allclose(lo, lon - dt * 2.5 / O.m_per_deg_lon(lat_mid), atol=1e-12)


def test_rk2_second_order_in_dt():
    """v = -a x (linear in lon): RK2 error falls ~4x when dt halves."""
    Z, Y, X = 3, 3, 201
    lon_ax = np.linspace(0.0, 2.0, X)
    uvw = np.zeros((3, Z, Y, X))
    aa = 1e-3
    win = O.Window(uvw, uvw, None, None, None, None, None, 0.0, 2.0, 0.0, 1.0, np.array([0.0, 1.0, 2.0]))
    # u such that dlon/dt = -a lon at lat 0: u = -a lon * m_per_deg_lon(0)
    uvw[0] = (-aa * lon_ax * O.m_per_deg_lon(np.zeros(1))[0])[None, None, :]
    errs = []
    for dt in (10.0, 5.0, 2.5, 1.25):
        lo = a(0.5)
        for _ in range(int(round(200.0 / dt))):
            lo, _, _ = O.advect_rk2(lo, a(0.0), a(1.0), win, 0.0, dt)
        errs.append(abs(lo[0] - 0.5 * math.exp(aa * 200.0)))
    for e1, e2 in zip(errs, errs[1:]):
        assert e1 / e2 > 3.5


def test_trilinear_reproduces_linear_field_exactly():
    Z, Y, X = 6, 5, 7
    lev = np.cumsum(np.r_[0.0, np.linspace(10, 400, Z - 1)])
    lon = np.linspace(-5.0, 1.0, X)
    lat = np.linspace(60.0, 58.0, Y)
    f = (1.0 + 0.3 * lon[None, None, :] - 0.7 * lat[None, :, None] + 0.002 * lev[:, None, None])[None]
    win = O.Window(*(None,) * 7, lon0=lon[0], lon1=lon[-1], lat0=lat[0], lat1=lat[-1], levels=lev)
    rng = np.random.default_rng(3)
    pl, pa, pz = rng.uniform(-5, 1, 50), rng.uniform(58, 60, 50), rng.uniform(0, lev[-1], 50)
    got = O.sample_stack(f, pl, pa, pz, win)[0]
    assert np.allclose(got, 1.0 + 0.3 * pl - 0.7 * pa + 0.002 * pz, atol=1e-9)


def test_zero_wind_langevin_gaussian_spread():
    """OU velocity integrated in position: variance matches the closed form
    (reference tests/test_physics.py:305-367, Philox noise)."""
    n, dt, tl, s2, steps = 30000, 1.0, 120.0, 2.0, 600
    w = np.zeros(n)
    z = np.full(n, 5000.0)
    uid = np.arange(n)
    for k in range(steps):
        eta = philox.normals(22, uid, k, 0)[:, 0]
        w = O.ou_update(w, tl, s2, dt, eta, 0.0)
        z = z + w * dt
    dz = z - 5000.0
    T = dt * steps
    exp_std = math.sqrt(2.0 * s2 * tl * (T - tl * (1.0 - math.exp(-T / tl))))
    std = dz.std(ddof=1)
    assert abs(dz.mean()) < 4.0 * std / math.sqrt(n)
    assert 0.85 * exp_std <= std <= 1.15 * exp_std
    c = (dz - dz.mean()) / std
    assert abs((c**3).mean()) < 0.2 and abs((c**4).mean() - 3.0) < 0.35


def test_well_mixed_drift_keeps_uniform_distribution():
    """Thomson drift keeps a uniform tracer uniform in sigma^2(z) linear; without
    it the low-sigma layer accumulates (reference tests/test_physics.py:370-444)."""
    H, n, tl, dt, steps = 1000.0, 20000, 100.0, 1.0, 3000
    s2b, s2t = 0.5, 2.0
    ds = (s2t - s2b) / H

    def evolve(drift_on):
        uid = np.arange(n)
        z = philox.uniforms(7, uid, 999999, 0)[:, 0] * H
        w = np.zeros(n)
        for k in range(steps):
            s2 = s2b + (s2t - s2b) * (z / H)
            dr = 0.5 * (1.0 + w**2 / s2) * ds if drift_on else 0.0
            w = O.ou_update(w, tl, s2, dt, philox.normals(7, uid, k, 0)[:, 0], dr * dt)
            z = z + w * dt
            below, above = z < 0.0, z > H
            z = np.where(below, -z, np.where(above, 2.0 * H - z, z))
            w = np.where(below | above, -w, w)
        return float((z < H / 3.0).mean())

    with_d, without = evolve(True), evolve(False)
    assert abs(with_d - 1.0 / 3.0) < 0.05
    assert without > with_d + 0.08


@pytest.mark.parametrize("z0,w0,zr,wr", [(200.0, 0.5, 200.0, 0.5), (0.0, -0.3, 0.0, -0.3),
                                          (-100.0, -1.2, 100.0, 1.2), (-1500.0, -2.0, 1500.0, 2.0)])
def test_reflection_known_answers(z0, w0, zr, wr):
    """Smooth-wall (z, w) -> (-z, -w) below the surface only (reference
    tests/test_physics.py:103-156), through one zero-noise substep."""
    z = a(z0)
    below = z < 0.0
    zz = np.where(below, 0.0 - z, z)
    ww = np.where(below, -a(w0), a(w0))
    assert zz[0] == zr and ww[0] == wr
