"""Pins the oracle's hourly resample and window stacks: the reference's
known-answer tests (tests/test_vertical_grid.py:38-135) re-expressed, plus
per-column numpy.interp on non-linear fields (a library routine)."""

from __future__ import annotations

import numpy as np

from oracle import resample as R
from glide_amd.synthetic import stretched_agl_levels


def desc_col(terrain_m):
    asl = np.array([12000.0, 6000.0, 3000.0, 1500.0, 800.0, 400.0, 150.0, 50.0])
    return asl - terrain_m


def regrid(field, h, targets):
    lower, frac, flipped = R.regrid_weights(h, targets)
    return R.apply_regrid(field, lower, frac, flipped)


def test_linear_field_recovered_exactly():
    h = desc_col(0.0)[:, None, None]
    f = (3.0 * h + 7.0)[None]
    t = np.array([50.0, 400.0, 1500.0, 6000.0])
    assert np.allclose(regrid(f, h, t)[0, :, 0, 0], 3.0 * t + 7.0, atol=1e-6)


def test_subsurface_levels_excluded():
    h = desc_col(1000.0)
    assert (h < 0).sum() >= 1
    f = np.where(h >= 0, np.arange(h.size) + 10.0, 999.0)[None, :, None, None]
    pos = np.where(h >= 0)[0]
    real_low = f[0, pos[np.argmin(h[pos])], 0, 0]
    out = regrid(f, h[:, None, None], np.array([0.0, 20.0, 40.0]))
    assert np.allclose(out[0, :, 0, 0], real_low) and not np.any(out == 999.0)


def test_constant_extrapolation_above_top_and_orientation():
    h = desc_col(0.0)
    f = np.arange(h.size, dtype=float)[None, :, None, None]
    assert regrid(f, h[:, None, None], np.array([50000.0]))[0, 0, 0, 0] == f[0, np.argmax(h), 0, 0]
    hd = desc_col(200.0)
    fd = (2.0 * hd + 1.0)[None, :, None, None]
    t = np.array([0.0, 100.0, 500.0, 2000.0])
    assert np.allclose(regrid(fd, hd[:, None, None], t), regrid(fd[:, ::-1], hd[::-1, None, None], t))


def test_nonlinear_field_matches_numpy_interp_per_column():
    rng = np.random.default_rng(5)
    Zp, Y, X = 12, 3, 4
    base = np.sort(rng.uniform(-300, 15000, Zp))[::-1].copy()
    h = base[:, None, None] + rng.normal(0, 20, (Zp, Y, X))
    h = np.sort(h, axis=0)[::-1].copy()
    f = rng.normal(size=(2, Zp, Y, X))
    t = np.array([0.0, 10.0, 35.0, 200.0, 1000.0, 7000.0, 20000.0])
    out = regrid(f, h, t)
    for y in range(Y):
        for x in range(X):
            col = h[::-1, y, x]
            keep = col >= 0  # sub-surface levels excluded
            for c in range(2):
                fc = f[c, ::-1, y, x]
                want = np.interp(t, col[keep], fc[keep])
                assert np.allclose(out[c, :, y, x], want, atol=1e-12)


def test_terrain_gradient_planar_slope_and_taper():
    lat = np.array([46.0, 45.0, 44.0])
    lon = np.array([0.0, 1.0, 2.0])
    terr = 100.0 * lon[None, :] * np.ones((3, 3))
    dhdx, dhdy = R.terrain_gradient(terr, lat, lon)
    assert np.allclose(dhdx, (100.0 / (111320.0 * np.cos(np.deg2rad(lat))))[:, None], rtol=1e-6)
    assert np.allclose(dhdy, 0.0, atol=1e-12)
    agl = np.array([0.0, 1000.0, 10000.0])
    w = np.zeros((3, 2, 2))
    out = R.slope_correct_w(w, np.ones_like(w), np.zeros_like(w), np.full((2, 2), 0.1), np.zeros((2, 2)),
                            agl, 10000.0)
    assert np.allclose(out[0], -0.1) and np.allclose(out[2], 0.0)


def test_stretched_levels_geometry():
    lev = stretched_agl_levels(60, 20000.0, 10.0)
    assert lev[0] == 0.0 and lev[-1] == 20000.0 and abs(lev[1] - 10.0) < 1e-9
    th = np.diff(lev)[:-1]
    r = th[1:] / th[:-1]
    assert np.allclose(r, r[0], rtol=1e-9) and r[0] > 1.0
    assert np.array_equal(lev, R.stretched_agl_levels(60, 20000.0, 10.0))


def test_source_unit_conversions_golden_values():
    """The reader's conversions (met_reader.py:1052-1078, 1268-1300, 1353-1385)
    on the reference's own vectors (tests/test_met_reader.py:509-617) and
    closed forms: a wrong sign, a missing de-accumulation or a dropped g
    fails."""
    from oracle import resample as OR

    # sensible heat flux: ECMWF downward-positive J/m^2 or W/m^2 -> upward-positive W/m^2
    assert np.allclose(OR.shf_to_upward(np.array([[-360_000.0]]), "accumulated"), 100.0)
    assert np.allclose(OR.shf_to_upward(np.array([[-100.0]]), "instantaneous"), 100.0)
    assert np.allclose(OR.shf_to_upward(np.array([[43_200.0]]), "accumulated"), -12.0)
    assert np.allclose(OR.shf_to_upward(np.array([[360_000.0]]), "accumulated"), -100.0)
    # omega -> w: subsidence-sign omega gives upward w; w = -omega / (rho g), rho = p / (R T)
    p = np.array([50_000.0, 85_000.0])
    t = np.full((2, 1, 1), 250.0, np.float32)
    w = OR.omega_to_w(np.full((2, 1, 1), -0.1, np.float32), t, p)
    rho = p / (287.05 * 250.0)
    np.testing.assert_allclose(w[:, 0, 0], 0.1 / (rho * 9.80665), rtol=1e-6)
    assert np.all(w > 0) and w.dtype == np.float32
    # geopotential -> height above ground: 9806.65 m^2/s^2 above a 1 km surface = 1 km AGL; below ground negative
    z = np.array([[[9806.65 * 2.0]], [[9806.65 * 0.5]]], np.float32)
    zs = np.array([[9806.65]], np.float32)
    h = OR.agl_from_geopotential(z, zs)
    np.testing.assert_allclose(h[:, 0, 0], [1000.0, -500.0], rtol=1e-6)
    assert h.dtype == np.float32


def test_synthetic_source_units_round_trip():
    """A met generated in source units converts back to the final-unit met
    of the same seed (to float32 rounding of the round trip)."""
    import dataclasses

    from glide_amd.synthetic import MetSpec, SyntheticMet
    from oracle import resample as OR

    spec = MetSpec(nx=16, ny=12, za=10, n_hours=2, seed=5)
    fin = SyntheticMet(spec).hour(1)
    raw = SyntheticMet(dataclasses.replace(spec, source_units=True)).hour(1)
    assert raw.units == "source" and raw.height_agl is None
    src3, src2, height = OR.convert_source_hour(raw)
    np.testing.assert_allclose(height, fin.height_agl, rtol=0, atol=2e-3)
    np.testing.assert_allclose(src3[2], fin.src3[2], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(src2[3], fin.src2[3], rtol=1e-6, atol=1e-4)
    np.testing.assert_array_equal(np.delete(src3, 2, 0), np.delete(fin.src3, 2, 0))
    assert not np.allclose(raw.src3[2], fin.src3[2])  # the raw w channel holds omega
