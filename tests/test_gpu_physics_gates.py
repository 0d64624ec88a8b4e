"""GPU: the reference's physics acceptance gates through the engine's fused
step (``glide_step`` / ``HannaScheme.step`` through the C ABI), with the
reference's set-ups, sizes and tolerances:

* OU autocorrelation and stationarity, solid-body rotation, the Taylor (1921)
  curve and its dt bias (reference tests/test_dispersion_analytic.py:38-243);
* the Langevin diffusion limit against Crank-Nicolson, with the teeth test
  that a halved or quartered near-surface K fails
  (tests/test_diffusion_pde.py:46-169);
* the analytic Gaussian-plume footprint (tests/test_plume_footprint.py:57-190);
* the well-mixed criterion through the production backward HannaScheme step,
  flat and density-weighted (tests/test_main_runtime.py:573-921).

The reference runs its engine-level gates by calling its OU, displacement,
reflection and RK2 primitives with prescribed (sigma_w, T_L); the fused step
has no separate primitives, so those gates run the whole step with
GLIDE_F_PRESCRIBED (sigma and T_L from the support record at every height; no
drift where sigma is uniform and d rho / dz is 0) or GLIDE_F_NO_TURB (the RK2
advection alone). Differences from the reference's set-ups: noise comes from
the engine's Philox stream (seeded with the reference's seeds); the step
integrates backward (z -= w dt, drift sign-flipped), which leaves the spread
statistics the same; in the plume the along-wind velocity u' is an OU process
with the same sigma as v' and w' (the prescribed mode has one sigma), where the
reference holds u' = 0: along-wind spreading of at most ~450 m over 4000 s
smooths the crosswind-integrated columns, which vary on kilometre scales, by
well under 1 %, but it also smears the front of the cloud, so the run goes on
500 s past the reference's 4000 s and the reference's bands (travel times up
to 4000 s) are checked on fully developed columns; the solid-body rotation runs near the pole, where
the step's metric 111320 max(|cos lat|, 0.05) is constant, so the linear
rotation field is a rigid rotation of the particles' metric coordinates."""

from __future__ import annotations

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

import physics_gates as G  # noqa: E402

from glide_amd import _lib as L  # noqa: E402
from glide_amd.footprint import FootprintGridder  # noqa: E402
from glide_amd.turbulence import GPUEngine, HannaScheme  # noqa: E402


def _homogeneous(sigma: float, t_l: float):
    return G.uniform_window(levels=[0.0, 20000.0], lon=np.linspace(0.0, 1.0, 4), lat=np.linspace(1.0, 0.0, 4),
                            sigma=sigma, t_l=t_l)


# --------------------------------------------------------------------------- #
# T5a: OU autocorrelation and stationarity
# --------------------------------------------------------------------------- #
def test_ou_autocorrelation_and_stationarity():
    n, sigma_w2, t_l, dt = 100_000, 0.7, 100.0, 5.0
    sigma = math.sqrt(sigma_w2)
    w0 = np.random.default_rng(4111).standard_normal(n) * sigma
    st = G.GpuStepper(_homogeneous(sigma, t_l), 0.5, 0.5, np.full(n, 10000.0), flags=L.F_PRESCRIBED, seed=4111,
                      w0=w0)
    for _ in range(int(2 * t_l / dt)):
        st.step(dt)
    n_frames = int(3 * t_l / dt) + 1
    frames = torch.empty((n_frames, n), dtype=torch.float64, device="cuda")
    for k in range(n_frames):
        frames[k] = st.state["w_prime"]
        st.step(dt)
    var = frames.var(unbiased=True).item()
    assert 0.97 * sigma_w2 <= var <= 1.03 * sigma_w2, var
    centered = frames - frames.mean()
    for tau_over_tl in (0.5, 1.0, 2.0):
        lag = int(round(tau_over_tl * t_l / dt))
        r = (centered[:-lag] * centered[lag:]).mean().item() / var
        assert abs(r - math.exp(-tau_over_tl)) < 0.02, (tau_over_tl, r)


# --------------------------------------------------------------------------- #
# T5b: solid-body rotation (RK2 advection)
# --------------------------------------------------------------------------- #
def test_solid_body_rotation_advection_returns_to_start():
    omega = 2.0 * math.pi / 3600.0
    period = 3600.0
    win = G.rotation_window(omega)
    lon0, lat0, xy0 = G.rotation_start()
    errors = []
    for dt in (60.0, 30.0, 15.0):
        st = G.GpuStepper(win, lon0, lat0, np.full(4, 500.0), flags=L.F_NO_TURB, seed=1)
        for _ in range(int(round(period / dt))):
            st.step(dt)
        P = st.P.cpu().numpy()
        errors.append(float(np.max(np.linalg.norm(G.rotation_xy(P[:, 0], P[:, 1]) - xy0, axis=1))))
    for i in range(len(errors) - 1):
        assert errors[i] / max(errors[i + 1], 1e-12) > 3.5, errors
    r_max = float(np.max(np.linalg.norm(xy0, axis=1)))
    assert errors[-1] < 1e-3 * r_max, errors
    # z and weight untouched by horizontal advection
    assert np.all(P[:, 2] == np.float32(500.0))
    assert np.all(P[:, 3] == np.float32(0.25))


# --------------------------------------------------------------------------- #
# T1: Taylor (1921) dispersion curve, and the position-integration dt bias
# --------------------------------------------------------------------------- #
def _spread(st, t_l, dt, checkpoints_tl):
    out, step = {}, 0
    for c in checkpoints_tl:
        target = int(round(c * t_l / dt))
        while step < target:
            st.step(dt)
            step += 1
        out[c] = torch.std(st.z.double(), unbiased=True).item()
    return out


def test_taylor_dispersion_curve_ballistic_to_diffusive():
    n, sigma_w2, t_l, dt = 100_000, 1.0, 100.0, 1.0
    sigma = math.sqrt(sigma_w2)
    w0 = np.random.default_rng(2201).standard_normal(n) * sigma
    st = G.GpuStepper(_homogeneous(sigma, t_l), 0.5, 0.5, np.full(n, 10000.0), flags=L.F_PRESCRIBED, seed=2201,
                      w0=w0)
    cps = [0.1, 0.3, 1.0, 3.0, 10.0, 30.0]
    sz = _spread(st, t_l, dt, cps)
    for c in cps:
        expected = math.sqrt(G.taylor_var(sigma_w2, t_l, c * t_l))
        assert abs(sz[c] - expected) / expected < 0.05, (c, sz[c], expected)
    t_bal = 0.1 * t_l
    assert abs(sz[0.1] - sigma * t_bal) / (sigma * t_bal) < 0.05
    k = sigma_w2 * t_l
    t_dif = 30.0 * t_l
    assert abs(sz[30.0] ** 2 - 2.0 * k * t_dif) / (2.0 * k * t_dif) < 0.08
    assert abs(st.P[:, 3].double().sum().item() - 1.0) < 1e-4  # weights untouched (float32 1/n each)


def test_taylor_dispersion_position_integration_bias_with_dt():
    t_l, sigma_w2 = 100.0, 1.0
    total_t = 5.0 * t_l
    expected = math.sqrt(G.taylor_var(sigma_w2, t_l, total_t))

    def spread_at_dt(dt, seed):
        n = 100_000
        w0 = np.random.default_rng(seed).standard_normal(n) * math.sqrt(sigma_w2)
        st = G.GpuStepper(_homogeneous(1.0, t_l), 0.5, 0.5, np.full(n, 10000.0), flags=L.F_PRESCRIBED, seed=seed,
                          w0=w0)
        for _ in range(int(round(total_t / dt))):
            st.step(dt)
        return torch.std(st.z.double(), unbiased=True).item()

    fine = spread_at_dt(1.0, 71)
    coarse = spread_at_dt(20.0, 71)
    assert abs(fine - expected) / expected < 0.02, fine
    assert abs(coarse - expected) / expected < 0.15, coarse


# --------------------------------------------------------------------------- #
# Diffusion limit vs Crank-Nicolson, and its teeth
# --------------------------------------------------------------------------- #
PDE_CHECKPOINTS = (300.0, 900.0, 1800.0)


@pytest.fixture(scope="module")
def diffusion():
    n = 100_000
    lev = G.pde_levels()
    win = G.uniform_window(levels=lev, lon=np.linspace(0.0, 1.0, 4), lat=np.linspace(1.0, 0.0, 4),
                           sigma=G.pde_sigma(lev), t_l=G.PDE_TL)
    rng = np.random.default_rng(3301)
    z0 = rng.random(n) * (G.PDE_Z0[1] - G.PDE_Z0[0]) + G.PDE_Z0[0]
    w0 = rng.standard_normal(n) * np.sqrt((G.PDE_K0 + G.PDE_A * np.minimum(z0, G.PDE_ZCAP)) / G.PDE_TL)
    st = G.GpuStepper(win, 0.5, 0.5, z0, flags=L.F_PRESCRIBED, seed=3301, w0=w0, z_ubl_m=0.0)
    ref, edges = G.pde_reference(PDE_CHECKPOINTS)
    out, t = {}, 0.0
    h_top = torch.tensor(G.PDE_HTOP, dtype=torch.float32, device="cuda")
    for tc in PDE_CHECKPOINTS:
        for _ in range(int(round((tc - t) / G.PDE_DT))):
            st.step(G.PDE_DT)
            z, w = st.P[:, 2], st.state["w_prime"]
            above = z > h_top
            z.copy_(torch.where(above, 2.0 * h_top - z, z))  # the reference's top reflection
            w.copy_(torch.where(above, -w, w))
        t = tc
        hist, _ = np.histogram(st.z.cpu().numpy(), bins=edges)
        out[tc] = dict(particles=hist / hist.sum(), **ref[tc])
    return out


def test_langevin_diffusion_limit_matches_pde(diffusion):
    for tc in PDE_CHECKPOINTS:
        got, want = diffusion[tc]["particles"], diffusion[tc]["true"]
        l1 = np.abs(got - want).sum()
        assert l1 < 0.08, (tc, l1)
        ns_got, ns_want = got[:3].sum(), want[:3].sum()
        assert abs(ns_got - ns_want) / ns_want < 0.12, (tc, ns_got, ns_want)


def test_diffusion_pde_discriminates_near_surface_k_errors(diffusion):
    for tc, min_half, min_quarter in ((300.0, 0.25, 0.45), (900.0, 0.20, 0.45), (1800.0, 0.12, 0.30)):
        got = diffusion[tc]["particles"]
        assert np.abs(got - diffusion[tc]["half"]).sum() > min_half, tc
        assert np.abs(got - diffusion[tc]["quarter"]).sum() > min_quarter, tc


# --------------------------------------------------------------------------- #
# Analytic Gaussian-plume footprint
# --------------------------------------------------------------------------- #
PLUME_STEPS = 800  # the reference's run: its bands end at 800 * 5 s = 4000 s of travel
PLUME_EXTRA = 100  # the front of the cloud, spread along-wind by u', has passed the last band column


@pytest.fixture(scope="module")
def plume():
    n = 200_000
    sigma = math.sqrt(G.PL_SIGMA2)
    rng = np.random.default_rng(9021)
    u0, v0, w0 = (rng.standard_normal(n) * sigma for _ in range(3))
    st = G.GpuStepper(G.plume_window(), 0.0, 0.0, np.full(n, G.PL_ZR), flags=L.F_PRESCRIBED, seed=9021, u0=u0,
                      v0=v0, w0=w0)
    grid = FootprintGridder(**G.plume_gridder_bounds(), n_time_bins=1, n_y=G.PL_NY, n_x=G.PL_NX, device="cuda",
                            dtype=torch.float32)
    t_idx = torch.zeros(n, dtype=torch.int64, device="cuda")
    active = torch.ones(n, dtype=torch.bool, device="cuda")
    for _ in range(PLUME_STEPS + PLUME_EXTRA):
        st.step(G.PL_DT)
        grid.accumulate(st.P[:, :3], active, st.P[:, 3], t_idx, t_idx, G.PL_DT)
    raw = grid.tensor.reshape(1, 1, 2, G.PL_NY, G.PL_NX)[0, 0, 0].double().cpu().numpy()
    return raw, G.plume_analytic(PLUME_STEPS)


def test_footprint_matches_analytic_gaussian_plume_and_magnitude_and_width(plume):
    raw, ref = plume
    t_l = G.PL_TL
    G.plume_checks(raw, ref, bands=[(2 * t_l, 5 * t_l, 0.05, 0.02), (5 * t_l, 10 * t_l, 0.15, 0.05),
                                    (10 * t_l, PLUME_STEPS * G.PL_DT, 0.15, 0.05)],
                   corr_min=0.995, ratio_tol=0.03, width_tol=0.05)


# --------------------------------------------------------------------------- #
# Well-mixed criterion through the production backward HannaScheme step
# --------------------------------------------------------------------------- #
def _run_wmc(met, z0, blh, seed, n_steps=1500, dt=15.0):
    scheme = HannaScheme(meander_enabled=False, seed=seed)
    engine = GPUEngine(device="cuda")
    n = z0.size
    particles = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    particles[:, 1] = 45.0
    particles[:, 2] = torch.from_numpy(z0.astype(np.float32)).cuda()
    particles[:, 3] = 1.0 / n
    state = scheme.initialize_state(n, device=torch.device("cuda"), dtype=torch.float32)
    active = torch.ones(n, dtype=torch.bool, device="cuda")
    for _ in range(n_steps):
        particles, state = scheme.step(particles, state, met, t_alpha=0.5, dt_seconds=dt, active_mask=active,
                                       engine=engine)
        above = particles[:, 2] > blh  # the reference's manual reflection at the BL top
        particles[:, 2] = torch.where(above, 2.0 * blh - particles[:, 2], particles[:, 2])
        state["w_prime"].copy_(torch.where(above, -state["w_prime"], state["w_prime"]))
    return particles[:, 2].cpu().numpy()


def test_v1_well_mixed_hanna_backward_path():
    blh = 3000.0
    met = G.wmc_met(blh_m=blh)
    z0 = np.random.default_rng(202605301).random(6000) * blh
    z = _run_wmc(met, z0, blh, seed=202605301)
    in_bl = (z >= 0.0) & (z <= blh)
    assert in_bl.mean() > 0.99
    counts, _ = np.histogram(z[in_bl], bins=np.linspace(0.0, blh, 11))
    interior = counts[2:8]
    rel_rms = float(np.sqrt(np.mean(((interior - interior.mean()) / interior.mean()) ** 2)))
    assert rel_rms < 0.15, (rel_rms, counts.tolist())


def test_v1_density_weighted_well_mixed():
    blh, z_top, p_bot, p_top, t_k = 5000.0, 6000.0, 1000.0, 500.0, 280.0
    met = G.wmc_met(blh_m=blh, p_top_hpa=p_top, z_top=z_top)
    rng = np.random.default_rng(202605302)
    n = 6000
    rho_max = float(G.density_at(0.0, p_bot_hpa=p_bot, p_top_hpa=p_top, z_top_m=z_top, t_kelvin=t_k))
    z_init: list[float] = []
    while len(z_init) < n:
        z_try = rng.uniform(0.0, blh, size=n - len(z_init))
        rho_try = G.density_at(z_try, p_bot_hpa=p_bot, p_top_hpa=p_top, z_top_m=z_top, t_kelvin=t_k)
        z_init.extend(z_try[rng.uniform(0.0, rho_max, size=z_try.shape) < rho_try].tolist())
    z = _run_wmc(met, np.array(z_init[:n]), blh, seed=202605302)
    in_bl = (z >= 0.0) & (z <= blh)
    assert in_bl.mean() > 0.99
    edges = np.linspace(0.0, blh, 11)
    counts, _ = np.histogram(z[in_bl], bins=edges)
    centres = 0.5 * (edges[:-1] + edges[1:])
    rho = G.density_at(centres, p_bot_hpa=p_bot, p_top_hpa=p_top, z_top_m=z_top, t_kelvin=t_k)
    expected = counts.sum() * rho / rho.sum()
    ic, ie = counts[3:7], expected[3:7]
    rel_rms = float(np.sqrt(np.mean(((ic - ie) / ie) ** 2)))
    flat_rms = float(np.sqrt(np.mean(((ic - ic.mean()) / ic.mean()) ** 2)))
    assert rel_rms < 0.15, (rel_rms, counts.tolist())
    assert flat_rms > 0.5 * rel_rms or flat_rms < 0.05
