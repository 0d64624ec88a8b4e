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
