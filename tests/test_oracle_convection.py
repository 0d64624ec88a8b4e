"""Pins the oracle's convection: thermodynamic known answers and matrix
invariants from the reference (tests/test_convection.py:63-323) and a
brute-force check of the backward sampling distribution."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import convection as CV


def moist_column():
    p = np.linspace(100000.0, 20000.0, 30)
    z = np.linspace(0.0, 11000.0, 30)
    t = 300.0 - 0.0065 * z
    q = 0.018 * np.exp(-z / 2500.0)
    return t, q, p


def test_thermodynamic_known_answers():
    e = CV.saturation_vapor_pressure(np.array([273.15, 283.15, 293.15, 303.15])) / 100.0
    ref = np.array([6.11, 12.27, 23.37, 42.43])
    assert (np.abs(e - ref) / ref).max() < 0.01
    q = np.array([1e-3, 5e-3, 1e-2, 2e-2])
    assert np.allclose(CV.e_to_q(CV.q_to_e(q, 1e5), 1e5), q, atol=1e-12)
    tl = CV.lcl_temperature_bolton(293.15, 0.008, 101325.0)
    pl = CV.lcl_pressure_poisson(293.15, tl, 101325.0)
    assert tl < 293.15 and 80000.0 < pl < 100000.0
    the = CV.equivalent_potential_temperature(np.array([290.0]), np.array([0.01]), np.array([95000.0]))
    t2 = CV.lift_parcel_moist(the, np.array([70000.0]))
    the2 = CV.equivalent_potential_temperature(t2, CV.saturation_specific_humidity(t2, 70000.0), np.array([70000.0]))
    assert abs(the2 - the)[0] / the[0] < 0.01


def test_matrix_zero_without_cloud_and_above_top():
    t, q, p = moist_column()
    m = CV.layer_masses(p)
    assert np.all(CV.mass_flux_matrix(m, -1, -1, 1.0) == 0.0)
    assert np.all(CV.mass_flux_matrix(m, 3, 3, 1.0) == 0.0)
    lcl, lnb, cape, dtv = CV.compute_lcl_lnb_cape(t, q, p)
    assert lcl >= 0 and lnb > lcl and cape > 0
    f = CV.mass_flux_matrix(m, lcl, lnb, 0.05 * 3600.0)
    for r in range(lnb + 1, f.shape[0]):
        assert np.all(f[r] == 0) and np.all(f[:, r] == 0)
    assert f[:, :lcl].sum() > 0


def test_matrix_non_divergent_and_mass_stationary_both_directions():
    t, q, p = moist_column()
    m = CV.layer_masses(p)
    lcl, lnb, _, _ = CV.compute_lcl_lnb_cape(t, q, p)
    f = CV.mass_flux_matrix(m, lcl, lnb, 0.05 * 3600.0)
    assert np.allclose(np.diag(f), 0.0)
    assert np.allclose(f.sum(1), f.sum(0), rtol=1e-9, atol=1e-12)
    ff = f + np.diag(m - f.sum(1))
    for P in (ff / m[:, None], ff.T / m[:, None]):
        assert np.allclose(P.sum(1), 1.0, atol=1e-9) and (P >= -1e-12).all()
        assert np.allclose(m @ P, m, rtol=1e-9)


def test_float32_matrix_rounds_the_float64_matrix():
    # The float32 restatement is the float64 matrix up to float32 rounding:
    # a dropped term (phi, the CFL scale) or a transposed product would show
    # at the 1e-2 level, rounding stays near 1e-6.
    t, q, p = moist_column()
    m = CV.layer_masses(p)
    lcl, lnb, _, _ = CV.compute_lcl_lnb_cape(t, q, p)
    for m_b in (0.05 * 3600.0, 5.0 * 3600.0):  # the second trips the CFL cap
        f64 = CV.mass_flux_matrix(m, lcl, lnb, m_b)
        f32 = CV.mass_flux_matrix_f32(m.astype(np.float32), lcl, lnb, m_b)
        assert f32.dtype == np.float32
        np.testing.assert_allclose(f32, f64, rtol=2e-5, atol=1e-6 * np.abs(f64).max())
        err = CV.column_sum_error(f32, m)
        assert 0.0 <= err < 1e-5
        assert CV.column_sum_error(f64, m) < 1e-12
    assert (f64.sum(1) / m).max() == pytest.approx(0.9, rel=1e-9)
    assert CV.mass_flux_matrix_f32(m.astype(np.float32), -1, -1, 1.0).sum() == 0.0


def test_backward_sampling_distribution_brute_force():
    t, q, p = moist_column()
    m = CV.layer_masses(p)
    lcl, lnb, _, _ = CV.compute_lcl_lnb_cape(t, q, p)
    f = CV.mass_flux_matrix(m, lcl, lnb, 0.2 * 3600.0)
    prob = CV.backward_probabilities(f, np.maximum(m, 1e-6))
    host = np.full(200000, lcl + 2)
    u = np.random.default_rng(0).random(host.size)
    dest = CV.sample_destinations(prob, host, u)
    freq = np.bincount(dest, minlength=m.size) / host.size
    want = prob[lcl + 2] / prob[lcl + 2].sum()
    assert np.abs(freq - want).max() < 0.005
    # nearest-level host: ties go up, below level 0 maps to 0
    h = np.array([0.0, 10.0, 30.0])
    assert list(CV.host_levels(np.array([-5.0, 4.9, 5.0, 25.0, 99.0]), h)) == [0, 0, 1, 2, 2]


def _moist_unstable_column(n_lev=12):
    """The reference's deep moist-unstable sounding (tests/test_convection.py:
    160-182): 300 K and 14 g/kg at 1000 hPa, 6.5 K/km to a 200 K floor."""
    p_hpa = np.exp(np.linspace(np.log(1000.0), np.log(100.0), n_lev))
    z_km = -8.0 * np.log(p_hpa / 1000.0)
    t = np.maximum(300.0 - 6.5 * z_km, 200.0)
    q = np.maximum(0.014 * np.exp(-(1000.0 - p_hpa) / 400.0), 1e-6)
    return t, q, p_hpa * 100.0


def _dry_stable_column(n_lev=8):
    """tests/test_convection.py:185-196: +2 K/km inversion, bone dry."""
    p_hpa = np.linspace(1000.0, 200.0, n_lev)
    z_km = -8.0 * np.log(p_hpa / 1000.0)
    return np.minimum(280.0 + 2.0 * z_km, 290.0), np.full(n_lev, 1e-5), p_hpa * 100.0


def test_lcl_lnb_cape_reference_soundings():
    """tests/test_convection.py:198-226 restated on the oracle."""
    lcl, lnb, cape, dtv = CV.compute_lcl_lnb_cape(*_moist_unstable_column())
    assert lcl > 0 and lnb > lcl and cape > 100.0 and dtv > 0.0
    _, _, cape, _ = CV.compute_lcl_lnb_cape(*_dry_stable_column())
    assert cape < 50.0


def test_lcl_lnb_cape_against_an_independent_solve():
    """The parcel path re-derived without the oracle's bisection: above the LCL
    the temperature is the root of theta_e(T, q_sat(T, p), p) = theta_e(surface)
    (scipy brentq to 1e-12 K), below it the dry adiabat; LCL = first level with
    p < p_LCL, LNB = first level above with buoyancy <= 0, CAPE = sum of
    g * max(buoyancy, 0) * dz (hypsometric dz) from the LCL up. The Bolton LCL
    itself is checked against saturation: the dry-lifted parcel at (T_LCL,
    p_LCL) holds q_sat within Bolton's ~0.1 K accuracy (a few % of q)."""
    from scipy.optimize import brentq

    for t, q, p in (_moist_unstable_column(), _moist_unstable_column(40), moist_column()):
        the = CV.equivalent_potential_temperature(t[0], q[0], p[0])
        tl = CV.lcl_temperature_bolton(t[0], q[0], p[0])
        pl = CV.lcl_pressure_poisson(t[0], tl, p[0])
        assert abs(CV.saturation_specific_humidity(tl, pl) / q[0] - 1.0) < 0.03
        tp = np.empty_like(t)
        for k in range(t.size):
            if p[k] >= pl:
                tp[k] = CV.potential_temperature(t[0], p[0]) * (p[k] / CV.P0) ** CV.KAPPA
            else:
                f = lambda T, k=k: CV.equivalent_potential_temperature(  # noqa: E731
                    T, CV.saturation_specific_humidity(T, p[k]), p[k]) - the
                # no root in the bracket (the adiabat leaves [180, 350] K aloft):
                # the bisection converges onto the bracket end. At 350 K and
                # low pressure q_sat is so large that theta_e's exponent
                # overflows to inf, which still orders the bracket correctly.
                with np.errstate(over="ignore"):
                    tp[k] = 180.0 if f(180.0) > 0 else (350.0 if f(350.0) <= 0 else
                                                        brentq(f, 180.0, 350.0, xtol=1e-12))
        qp = np.where(p >= pl, q[0], CV.saturation_specific_humidity(tp, p))
        tvp, tve = CV.virtual_temperature(tp, qp), CV.virtual_temperature(t, q)
        b = (tvp - tve) / tve
        i_lcl = int(np.argmax(p < pl))
        i_lnb = next((k for k in range(i_lcl + 1, t.size) if b[k] <= 0), t.size - 1)
        cape = sum(CV.GRAVITY * max(b[k], 0.0) * CV.R_DRY * 0.5 * (t[k - 1] + t[k]) * np.log(p[k - 1] / p[k])
                   / CV.GRAVITY for k in range(max(i_lcl, 1), t.size))
        lcl, lnb, cape_o, dtv = CV.compute_lcl_lnb_cape(t, q, p)
        assert (lcl, lnb) == (i_lcl, i_lnb)
        assert cape_o == pytest.approx(cape, rel=1e-6, abs=1e-9)
        assert dtv == pytest.approx(tvp[i_lcl + 1] - tve[i_lcl + 1], abs=1e-6)


def test_column_cape_lerps_and_clamps_per_column():
    t1, q1, p = _moist_unstable_column(12)
    t2, q2, _ = _dry_stable_column(12)
    t2 = np.minimum(280.0 + 2.0 * (-8.0 * np.log(p / 1e5)), 290.0)
    T = np.stack([t1, t2], -1)[:, None, :].astype(np.float32)  # [Z, 1, 2]
    Q = np.stack([q1, np.full(12, -1e-4)], -1)[:, None, :].astype(np.float32)
    for al in (0.0, 0.3, 1.0):
        lcl, lnb, cape, dtv = CV.column_cape(T, T + np.float32(1.0), Q, Q, al, p / 100.0)
        for c in range(2):
            ts, te = T[:, 0, c].astype(np.float64), (T[:, 0, c] + np.float32(1.0)).astype(np.float64)
            want = CV.compute_lcl_lnb_cape(ts * (1.0 - al) + te * al, np.maximum(Q[:, 0, c].astype(np.float64), 0.0), p)
            got = (lcl[0, c], lnb[0, c], cape[0, c], dtv[0, c])
            assert got[:2] == want[:2]
            assert got[2] == pytest.approx(want[2], rel=1e-12, abs=1e-12)
            assert got[3] == pytest.approx(want[3], rel=1e-12, abs=1e-12)
    assert cape[0, 1] < 50.0 and lcl[0, 0] >= 0 and cape[0, 0] > 100.0  # the dry inversion column stays stable
    lev = np.arange(12, dtype=np.float32) * 100.0
    conv = CV.conv_descriptors_from_lift(lcl, lnb, cape, dtv, lev)
    assert conv.shape == (4, 1, 2) and conv[1, 0, 0] == lev[lcl[0, 0]] and conv[2, 0, 0] == lev[lnb[0, 0]]
    none = CV.conv_descriptors_from_lift(np.array([[-1]]), np.array([[-1]]), np.zeros((1, 1)), -np.ones((1, 1)), lev)
    assert np.isinf(none[1, 0, 0]) and np.isinf(none[2, 0, 0])
