"""GPU: the reference's Hanna (1982) / FLEXPART unit tests (tests/test_hanna.py
:48-364 of the reference), run against the engine's own device code through
``glide_amd.hanna`` (``glide_hanna_eval`` / ``glide_windowed_std``): the
float32 functions the fused step kernel and the window build execute.

Same setups and assertions as the reference. The reference evaluates in
float64 and some of its bounds are below float32 resolution: those bounds are
restated as float32 ones (a few units in the last place, relative), each
marked where it appears."""

from __future__ import annotations

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd.hanna import (  # noqa: E402
    EARTH_ROTATION_RATE_S,
    FT_KZ_CEIL_M2_S,
    FT_KZ_FLOOR_M2_S,
    air_density,
    brunt_vaisala_squared,
    convective_velocity,
    coriolis_parameter,
    free_trop_diffusivity,
    free_trop_sigma_TL,
    gradient_richardson,
    in_bl_sigma_TL,
    obukhov_length,
    potential_temperature,
    surface_layer_sigma_w,
)
from glide_amd.turbulence import HannaScheme  # noqa: E402

F32 = 4e-7  # float32 relative resolution for the reference's float64 bounds (a few ulp)


def T(*v):
    return torch.tensor(v, dtype=torch.float64, device="cuda")


def close(got, want, rel=F32, abs_=0.0):
    return abs(float(got) - want) <= rel * abs(want) + abs_


def test_coriolis_parameter_signs_and_magnitude():
    f = coriolis_parameter(T(0.0, 30.0, 45.0, 60.0, -45.0)).cpu()
    assert abs(f[0].item()) < 1e-10
    expected_45 = 2.0 * EARTH_ROTATION_RATE_S * math.sin(math.radians(45.0))
    assert close(f[2].item(), expected_45)  # reference: 1e-12 absolute (float64)
    assert f[1].item() > 0
    assert f[3].item() > f[2].item() > f[1].item() > 0
    assert f[4].item() < 0
    assert abs(f[4].item() + f[2].item()) <= 2 * F32 * abs(f[2].item())  # reference: 1e-12


def test_air_density_standard_conditions():
    rho = air_density(T(101325.0), T(288.0))
    assert abs(rho.item() - 1.225) < 0.01


def test_obukhov_length_signs():
    L = obukhov_length(T(0.4, 0.4, 0.4), T(100.0, -50.0, 0.0), T(288.0, 288.0, 288.0),
                       T(101325.0, 101325.0, 101325.0)).cpu()
    assert L[0].item() < 0
    assert L[1].item() > 0
    assert math.isinf(L[2].item())


def test_convective_velocity_zero_when_stable_or_neutral():
    w_star = convective_velocity(T(1000.0, 1000.0, 1000.0), T(200.0, -50.0, 0.0), T(288.0, 288.0, 288.0),
                                 T(101325.0, 101325.0, 101325.0)).cpu()
    assert w_star[0].item() > 0
    assert w_star[1].item() == 0.0
    assert w_star[2].item() == 0.0
    rho = 101325.0 / (287.05 * 288.0)
    expected_w_star = ((9.80665 * 1000.0 * 200.0) / (288.0 * rho * 1005.0)) ** (1.0 / 3.0)
    assert close(w_star[0].item(), expected_w_star, rel=2 * F32)  # reference: 1e-6 absolute


def test_in_bl_sigma_at_z_zero_matches_ustar_scaling():
    su, sv, sw, *_ = (t.cpu() for t in in_bl_sigma_TL(
        T(0.0, 0.0, 0.0), T(1000.0, 1000.0, 1000.0), T(0.5, 0.5, 0.5), T(0.0, 0.0, 1.5), T(5.0, 0.0, -5.0),
        T(45.0, 45.0, 45.0)))
    for i in (0, 1):  # stable, neutral: sigma_u = 2.0 u*, sigma_v = sigma_w = 1.3 u*
        assert close(su[i].item(), 2.0 * 0.5)  # reference: 1e-9 absolute
        assert close(sw[i].item(), 1.3 * 0.5)
        assert close(sv[i].item(), 1.3 * 0.5)
    assert close(sw[2].item(), math.sqrt(1.8 * 0.5**2))  # reference: 1e-6 absolute
    assert not torch.allclose(su, sv)


def test_in_bl_stable_sigma_vanishes_at_top_of_bl():
    _, _, sw, *_ = in_bl_sigma_TL(T(1000.0), T(1000.0), T(0.5), T(0.0), T(5.0), T(45.0))
    assert sw.item() < 1e-2


def test_in_bl_stable_timescales_match_hanna():
    ustar_val, blh_val = 0.3, 500.0
    heights = [25.0, 100.0, 250.0, 400.0]
    z = T(*heights)
    out = [t.cpu() for t in in_bl_sigma_TL(z, torch.full_like(z, blh_val), torch.full_like(z, ustar_val),
                                           torch.zeros_like(z), torch.full_like(z, 5.0),
                                           torch.full_like(z, 45.0))]
    sigma_u, sigma_v, sigma_w, t_lu, t_lv, t_lw = out
    for i, z_val in enumerate(heights):
        zeta = z_val / blh_val
        su = 2.0 * ustar_val * (1.0 - zeta)
        sv = 1.3 * ustar_val * (1.0 - zeta)
        assert close(sigma_u[i].item(), su)  # reference: 1e-9 absolute
        assert close(sigma_v[i].item(), sv)
        assert close(sigma_w[i].item(), sv)
        assert close(t_lu[i].item(), 0.15 * blh_val / su * zeta**0.5, rel=4 * F32)  # reference: 1e-6 absolute
        assert close(t_lv[i].item(), 0.07 * blh_val / sv * zeta**0.5, rel=4 * F32)
        assert close(t_lw[i].item(), 0.10 * blh_val / sv * zeta**0.8, rel=8 * F32)  # zeta^0.8 via exp2/log2
    assert not torch.allclose(t_lv, 0.467 * t_lu, rtol=1e-3)
    assert torch.allclose(t_lv / t_lu, torch.full_like(t_lu, 0.718), rtol=1e-3)


def test_in_bl_regime_selection_at_boundaries():
    z, blh, ustar = T(100.0), T(1000.0), T(0.5)
    lat = T(45.0)
    su_neu = in_bl_sigma_TL(z, blh, ustar, T(0.0), T(0.5), lat)[0].item()
    su_stab = in_bl_sigma_TL(z, blh, ustar, T(0.0), T(1.5), lat)[0].item()
    su_unst = in_bl_sigma_TL(z, blh, ustar, T(1.5), T(-1.5), lat)[0].item()
    su_deep = in_bl_sigma_TL(z, blh, ustar, T(1.5), T(-10.0), lat)[0].item()
    assert su_deep > su_unst > 0
    assert su_neu != su_stab


def test_surface_layer_sigma_w_regimes():
    sw = surface_layer_sigma_w(T(10.0, 10.0, 10.0), T(0.5, 0.5, 0.5), T(float("inf"), -100.0, 200.0)).cpu()
    assert abs(sw[0].item() - 0.65) < 1e-6
    assert abs(sw[1].item() - 0.65 * (1.3 ** (1.0 / 3.0))) < 1e-6
    assert abs(sw[2].item() - 0.65) < 1e-6


def test_surface_layer_sigma_w_stable_is_constant_in_height():
    sw = surface_layer_sigma_w(T(10.0, 100.0), T(0.5, 0.5), T(200.0, 5.0)).cpu()
    assert close(sw[0].item(), 0.65)  # reference: 1e-9 absolute (0.65 is not a float32)
    assert sw[0].item() == sw[1].item()


def test_potential_temperature_reference_level():
    kappa = 287.05 / 1005.0
    theta = potential_temperature(T(280.0, 250.0), T(100000.0, 50000.0)).cpu()
    assert abs(theta[0].item() - 280.0) < 1e-9
    assert theta[1].item() > 250.0
    assert close(theta[1].item(), 250.0 * 2.0**kappa, rel=4 * F32)  # reference: 1e-6 absolute


def test_brunt_vaisala_sign_follows_stratification():
    theta = T(300.0, 300.0)
    assert brunt_vaisala_squared(theta, T(0.005, 0.005))[0].item() > 0
    assert brunt_vaisala_squared(theta, T(-0.005, -0.005))[0].item() < 0


def test_gradient_richardson_and_diffusivity_shutoff():
    z = T(1000.0, 1000.0, 1000.0)
    shear = T(0.01, 0.01, 0.01)
    ri = gradient_richardson(T(0.0, 1e-5, 1e-3), shear.pow(2))
    k_z = free_trop_diffusivity(z, shear, ri).cpu()
    assert k_z[0].item() > k_z[1].item() >= k_z[2].item()
    assert close(k_z[2].item(), FT_KZ_FLOOR_M2_S)  # reference: 1e-9 absolute (0.1 is not a float32)
    assert k_z[0].item() <= FT_KZ_CEIL_M2_S


def test_free_trop_sigma_TL_satisfies_diffusivity_identity():
    k_z = T(0.1, 1.0, 10.0)
    sigma_w, t_lw = free_trop_sigma_TL(k_z, T(1e-4, 1e-4, 1e-4))
    assert torch.allclose(sigma_w.double().pow(2) * t_lw.double(), k_z, rtol=1e-5)


def test_free_trop_diffusivity_never_zero():
    k_z = free_trop_diffusivity(T(2000.0), T(0.001), T(100.0))
    assert k_z.item() >= np.float32(FT_KZ_FLOOR_M2_S)


def test_meander_windowed_std_uniform_field_is_zero():
    field = torch.full((3, 5, 5), 7.0, dtype=torch.float64, device="cuda")
    std = HannaScheme._windowed_std(field, radius=1)
    assert std.shape == field.shape
    assert torch.allclose(std, torch.zeros_like(std), atol=1e-6)


@pytest.mark.parametrize("radius", [1, 2])
def test_meander_windowed_std_matches_numpy_population_std(radius):
    rng = np.random.default_rng(0)
    arr = rng.standard_normal((1, 6, 7))
    got = HannaScheme._windowed_std(torch.as_tensor(arr, device="cuda"), radius=radius).cpu().numpy()
    z, ny, nx = arr.shape
    want = np.empty_like(arr)
    for k in range(z):
        for j in range(ny):
            for i in range(nx):
                window = arr[k, max(0, j - radius): j + radius + 1, max(0, i - radius): i + radius + 1]
                want[k, j, i] = window.std()
    np.testing.assert_allclose(got, want, atol=2e-6)  # reference: 1e-10 (float64)
